"""Stand-in for clawpack.riemann.euler_4wave_2D_constants: component indices."""
density = 0
x_momentum = 1
y_momentum = 2
energy = 3
num_eqn = 4
num_waves = 4

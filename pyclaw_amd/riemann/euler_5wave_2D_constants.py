"""Stand-in for clawpack.riemann.euler_5wave_2D_constants: component indices."""
density = 0
x_momentum = 1
y_momentum = 2
energy = 3
tracer = 4
num_eqn = 5
num_waves = 5

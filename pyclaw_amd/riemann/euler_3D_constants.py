"""Stand-in for clawpack.riemann.euler_3D_constants: component indices."""
density = 0
x_momentum = 1
y_momentum = 2
z_momentum = 3
energy = 4
num_eqn = 5
num_waves = 3

"""Stand-in for clawpack.riemann.shallow_roe_with_efix_2D_constants: component indices."""
depth = 0
x_momentum = 1
y_momentum = 2
num_eqn = 3
num_waves = 3

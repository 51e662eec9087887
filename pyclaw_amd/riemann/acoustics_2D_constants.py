"""Stand-in for clawpack.riemann.acoustics_2D_constants: component indices."""
pressure = 0
x_velocity = 1
y_velocity = 2
num_eqn = 3
num_waves = 2

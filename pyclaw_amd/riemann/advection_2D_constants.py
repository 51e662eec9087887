"""Stand-in for clawpack.riemann.advection_2D_constants: component indices."""
q = 0
num_eqn = 1
num_waves = 1

"""Shallow water over bathymetry in f-wave form: the device form of
clawpack.riemann's shallow_bathymetry_fwave_2D (rpn2), as examples/shallow_2d/
sill.py uses it (problem_data 'grav', 'dry_tolerance', 'sea_level'; the
bathymetry in aux[0]).  The source -g h b_x is part of the f-wave
decomposition, so solver.fwave is True.  No transverse solver is paired with
it: dimension-split steps, or unsplit with transverse_waves = 0."""
from ._rp import RPHandle

kind = 11
num_eqn = 3
num_waves = 3
num_aux = 1
fwave = True
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

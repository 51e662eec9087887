"""Shallow water with the HLLE approximate Riemann solver: the device form of
clawpack.riemann's shallow_hlle_2D (rpn2), as examples/shallow_2d's radial
dam break uses it (problem_data['grav']).  No transverse solver is paired with
it: unsplit runs need transverse_waves = 0."""
from ._rp import RPHandle

kind = 9
num_eqn = 3
num_waves = 2
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

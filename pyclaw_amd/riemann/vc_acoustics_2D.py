"""Variable-coefficient acoustics, q = (p, u, v), with the density and sound
speed of each cell in aux[0], aux[1]: the device form of clawpack.riemann's
vc_acoustics_2D (rpn2 and rpt2), as examples/acoustics_2d_variable uses it."""
from ._rp import RPHandle

kind = 8
num_eqn = 3
num_waves = 2
num_aux = 2
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

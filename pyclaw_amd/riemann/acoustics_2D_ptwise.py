"""Stand-in for clawpack.riemann.acoustics_2D_ptwise: the pointwise form of
acoustics_2D (the same waves, speeds and transverse split, written per
interface).  The device solvers are pointwise already, so it is the same
device solver as acoustics_2D."""
from ._rp import RPHandle

kind = 1
num_eqn = 3
num_waves = 2
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

"""Stand-in for clawpack.riemann.advection_2D: metadata and the device solver tag."""
from ._rp import RPHandle

kind = 4
num_eqn = 1
num_waves = 1
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")
rpn3 = RPHandle(kind, "rpn3")

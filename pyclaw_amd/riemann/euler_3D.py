"""Stand-in for clawpack.riemann.euler_3D: metadata and the device solver tag."""
from ._rp import RPHandle

kind = 5
num_eqn = 5
num_waves = 3
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")
rpn3 = RPHandle(kind, "rpn3")

"""Stand-in for clawpack.riemann.euler_4wave_2D: metadata and the device solver tag."""
from ._rp import RPHandle

kind = 2
num_eqn = 4
num_waves = 4
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")
rpn3 = RPHandle(kind, "rpn3")

"""Stand-in for clawpack.riemann.shallow_roe_with_efix_2D: metadata and the device solver tag."""
from ._rp import RPHandle

kind = 3
num_eqn = 3
num_waves = 3
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")
rpn3 = RPHandle(kind, "rpn3")

"""Variable-coefficient advection of any number of fields with edge velocities
in two aux fields: the device form of the Riemann solver of
examples/advection_reaction_2d/rpn2_vc_advection_vector.f90 (rpn2 only; use it
with dimensional splitting, as that example does).  aux[0] = u at a cell's left
edge, aux[1] = v at its lower edge.  num_eqn is set on the solver."""
from ._rp import RPHandle

kind = 6
num_eqn = None
num_waves = 1
num_aux = 2
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

"""Scalar variable-coefficient advection q_t + u q_x + v q_y = 0 with the edge
velocities in aux (aux[0] = u at a cell's left edge, aux[1] = v at its lower
edge): the device form of clawpack.riemann's vc_advection_2D (rpn2 and the
transverse rpt2, which splits asdq with the transverse edge velocity of the
receiving cell), as examples/advection_2d_annulus uses it on a mapped grid with
a capacity function in aux[2]."""
from ._rp import RPHandle

kind = 7
num_eqn = 1
num_waves = 1
num_aux = 2
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

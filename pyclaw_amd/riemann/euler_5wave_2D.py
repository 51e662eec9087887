"""Euler equations with a passive tracer: the device form of clawpack.riemann's
euler_5wave_2D (Roe solver with the Harten-Hyman entropy fix; the tracer q[4] is
carried by a fifth wave, development/rp_approaches/rpn2_euler_5wave.f and
rpt2_euler_5wave.f), as examples/euler_2d/shock_bubble_interaction.py uses it
(problem_data['gamma'])."""
from ._rp import RPHandle

kind = 10
num_eqn = 5
num_waves = 5
fwave = False
rpn2 = RPHandle(kind, "rpn2")
rpt2 = RPHandle(kind, "rpt2")

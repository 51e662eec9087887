"""Limiter identifiers (pyclaw/limiters/tvd.py:74-77) and philim's Beam-Warming
case (classic/philim.f90:32-33)."""
minmod = 1
superbee = 2
vanleer = 3
MC = 4
beam_warming = 5

"""pyclaw_amd -- PyClaw's time-stepping hot path on the MI355X (gfx950).

The public names mirror clawpack.pyclaw (pyclaw/__init__.py:20-43) so that set-up
scripts run unchanged with `import pyclaw_amd as pyclaw` and
`from pyclaw_amd import riemann`.  The steps run in hand-written HIP kernels
(libclawhip.so); there is no CPU fallback.
"""
from . import limiters, riemann  # noqa: F401
from .cfl import CFL  # noqa: F401
from .classic import ClawSolver2D, ClawSolver3D  # noqa: F401
from .controller import Controller  # noqa: F401
from .geometry import Dimension, Domain, Patch  # noqa: F401
from .sharpclaw import SharpClawSolver2D  # noqa: F401
from .limiters import tvd  # noqa: F401
from .solution import Solution  # noqa: F401
from .solver import BC, CFLError  # noqa: F401
from .state import State  # noqa: F401

__all__ = ["Controller", "Dimension", "Patch", "Domain", "Solution", "State", "CFL",
           "ClawSolver2D", "ClawSolver3D", "SharpClawSolver2D", "BC", "limiters", "riemann"]

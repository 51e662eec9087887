"""Stand-ins for the clawpack.riemann modules the hot path supports.

The reference passes a Riemann-solver *module* to the solver constructor
(e.g. ClawSolver2D(riemann.euler_4wave_2D), examples/euler_2d/quadrants.py:80)
and reads its metadata from riemann.static (pyclaw/solver.py:230-238).  A GPU
kernel cannot call a host Fortran function pointer, so each module here names the
device solver the kernels dispatch to (`kind`); `rpn2._cpointer` carries the
same tag for callers that use the fmod calling convention.
"""
from . import static  # noqa: F401
from . import acoustics_2D, acoustics_2D_ptwise, advection_2D, euler_3D, euler_4wave_2D, shallow_roe_with_efix_2D  # noqa: F401
from . import euler_5wave_2D, euler_5wave_2D_constants  # noqa: F401
from . import shallow_hlle_2D, vc_acoustics_2D, vc_advection_2D, vc_advection_vector_2D  # noqa: F401
from . import shallow_bathymetry_fwave_2D  # noqa: F401

"""Frame files (pyclaw/fileio): the classic Clawpack ascii format."""
from . import ascii  # noqa: F401

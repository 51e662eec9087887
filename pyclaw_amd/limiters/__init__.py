from . import tvd  # noqa: F401

__all__ = ["tvd"]

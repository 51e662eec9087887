"""Shared helper for the Riemann-solver stand-in modules."""


class RPHandle(object):
    """Stands where the reference has an f2py function object (rp.rpn2); its
    `_cpointer` is the solver tag the fmod drop-in modules dispatch on."""

    def __init__(self, kind, name):
        self._cpointer = kind
        self.__name__ = name

    def __call__(self, *args, **kwargs):
        raise TypeError("device Riemann solvers run inside the HIP kernels; they are not callable on the host")

"""CFL holder (pyclaw/cfl.py:5-29).  The kernels reduce the Courant number on the
device (FP64 atomic max); the solver reads it back into this object."""


class CFL(object):
    def __init__(self, global_max):
        self._global_max = global_max

    def get_global_max(self):
        return self._global_max

    def get_cached_max(self):
        return self._global_max

    def set_local_max(self, new_local_max):
        self._global_max = new_local_max

    def set_global_max(self, new_global_max):
        self._global_max = new_global_max

    def update_global_max(self, new_local_max):
        self._global_max = max(self._global_max, new_local_max)

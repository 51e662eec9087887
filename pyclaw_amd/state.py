"""State (pyclaw/state.py:7-352) with a device-resident copy of q.

`state.q` is the reference's host array (meqn, mx, my[, mz]) in Fortran order;
users read and write it as before (e.g. `state.q[0,:,:] = ...` in qinit).  The
solvers keep the authoritative copy on the GPU between steps: reading `state.q`
downloads it, and a read or write marks the device copy stale so the next step
uploads it again.  Nothing is copied while a solver steps.
"""
import copy

import numpy as np

from .geometry import Domain, Patch


class State(object):
    def __init__(self, geom, num_eqn, num_aux=0):
        if isinstance(geom, Patch):
            self.patch = geom
        elif isinstance(geom, Domain):
            self.patch = geom.patches[0]
        else:
            raise Exception("A State object must be initialized with a Patch or Domain object.")
        self.problem_data = {}
        self.t = 0.
        self.index_capa = -1
        self.keep_gauges = False
        self.gauge_data = []
        self.p = None
        self.F = None
        self._q = self.new_array(num_eqn)
        self.aux = self.new_array(num_aux) if num_aux > 0 else None
        self._dev = None          # (q_tensor, comp_tensor_or_None) when the device is newer
        self._dev_stale = True    # device copy must be re-uploaded before use
        self._aux_dev = None      # device copy of aux (uploaded at every step that uses it)

    # ---- geometry passthrough
    @property
    def grid(self):
        return self.patch.grid

    def __getattr__(self, key):
        if key in ("num_dim", "p_centers", "p_nodes", "c_centers", "c_nodes"):
            return getattr(self.__dict__["patch"].grid, key)
        raise AttributeError(key)

    @property
    def num_eqn(self):
        if self._q is None:
            raise Exception("state.num_eqn has not been set.")
        return self._q.shape[0]

    @property
    def num_aux(self):
        return self.aux.shape[0] if self.aux is not None else 0

    # ---- q with lazy device synchronisation
    @property
    def q(self):
        self._pull()
        self._dev_stale = True
        return self._q

    @q.setter
    def q(self, value):
        self._pull_discard()
        self._q = value
        self._dev_stale = True

    def _pull(self):
        if self._dev is not None:
            from .device import q_from_device
            self._q = q_from_device(self._dev[0])
            self._dev = None

    def _pull_discard(self):
        self._dev = None

    def device_q(self, dtype_code, ensure=True):
        """(q, comp) device tensors for the solver; uploads when stale."""
        import torch
        from .device import q_to_device, torch_dtype
        want = torch_dtype(dtype_code)
        if self._dev is not None and self._dev[0].dtype == want and not self._dev_stale:
            return self._dev
        if self._dev is not None:
            self._pull()
        q = q_to_device(np.asarray(self._q), dtype_code)
        comp = torch.zeros_like(q) if dtype_code == 1 else None
        self._dev = (q, comp)
        self._dev_stale = False
        return self._dev

    def device_aux(self, dtype_code):
        """Planar device copy of aux in the state's arithmetic type.  aux is a
        plain host array that users (and before_step callbacks) change in place,
        so every call uploads it again."""
        from .device import q_to_device
        self._aux_dev = q_to_device(np.asarray(self.aux), dtype_code)
        return self._aux_dev

    def set_device_q(self, q, comp=None):
        """The solver hands back the newest state (device tensors)."""
        self._dev = (q, comp)
        self._dev_stale = False

    def new_array(self, dof):
        if dof == 0:
            return None
        shape = [dof] + list(self.grid.num_cells)
        return np.empty(shape, order="F")

    def get_q_global(self):
        return self.q.copy()

    def get_aux_global(self):
        return None if self.aux is None else self.aux.copy()

    def is_valid(self):
        valid = True
        if self.q is None or not self.q.flags["F_CONTIGUOUS"]:
            valid = False
        return valid

    def set_num_ghost(self, num_ghost):
        pass

    def set_cparam(self, fortran_module):
        """state.py:206-231: copy problem_data into module.cparam when it has one."""
        if hasattr(fortran_module, "cparam"):
            for key, value in self.problem_data.items():
                setattr(fortran_module.cparam, key, value)

    # ---- reference qbc helpers (state.py:239-303), used by the fmod drop-in
    def get_qbc_from_q(self, num_ghost, qbc):
        sl = (slice(None),) + tuple(slice(num_ghost, -num_ghost) for _ in range(self.num_dim))
        qbc[sl] = self.q
        return qbc

    def set_q_from_qbc(self, num_ghost, qbc):
        sl = (slice(None),) + tuple(slice(num_ghost, -num_ghost) for _ in range(self.num_dim))
        self.q = qbc[sl]

    def __copy__(self):
        return self.__deepcopy__()

    def __deepcopy__(self, memo={}):
        result = self.__class__(copy.deepcopy(self.patch), self.num_eqn, self.num_aux)
        result.__init__(copy.deepcopy(self.patch), self.num_eqn, self.num_aux)
        result._q = np.array(self.q, order="F", copy=True)
        if self.aux is not None:
            result.aux = copy.deepcopy(self.aux)
        result.t = self.t
        result.problem_data = copy.deepcopy(self.problem_data)
        result.index_capa = self.index_capa
        return result

    def __str__(self):
        return "PyClaw State object\nPatch dimensions: %s\nTime t=%s\nNumber of conserved quantities: %s\n" % (
            str(self.patch.grid.num_cells), self.t, self.num_eqn)

"""State (pyclaw/state.py:7-352) with a device-resident copy of q.

`state.q` is the reference's host array (meqn, mx, my[, mz]) in Fortran order;
users read and write it as before (e.g. `state.q[0,:,:] = ...` in qinit).  The
solvers keep the authoritative copy on the GPU between steps: reading `state.q`
downloads it, and the next step uploads it again only if the host array was
changed meanwhile (into the same device buffers).  Nothing is copied while a
solver steps.
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
        self._dev = None          # (q_tensor, comp_tensor_or_None) on the device
        self._dev_stale = True    # the host array was handed out since the device copy
        self._host_current = True # the host array holds the newest state
        self._snap = None         # what the last download gave (see device_q)
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
        """Download the device state into the host array once per step run; a
        copy of what was downloaded tells device_q later whether the host array
        was changed meanwhile."""
        if self._dev is not None and not self._host_current:
            from .device import q_from_device
            self._q = q_from_device(self._dev[0])
            self._snap = self._q.copy(order="K")
            self._host_current = True

    def _pull_discard(self):
        self._dev = None
        self._snap = None
        self._host_current = True

    def device_q(self, dtype_code, ensure=True):
        """(q, comp) device tensors for the solver.  After the host array was
        handed out (state.q), it is compared with what was downloaded: unchanged,
        the device tensors (and the Kahan compensation) stay as they are;
        changed, the new values go into the same device buffers (captured graphs
        keep their addresses) and the compensation restarts at zero."""
        import torch
        from .device import q_to_device, torch_dtype
        want = torch_dtype(dtype_code)
        dev = self._dev
        if dev is not None and dev[0].dtype == want:
            if not self._dev_stale:
                return dev
            if (self._snap is not None and self._q is not None
                    and self._q.shape == self._snap.shape
                    and np.array_equal(self._q, self._snap)):
                self._dev_stale = False
                return dev
        if dev is not None and not self._host_current:
            self._pull()
        q = q_to_device(np.asarray(self._q), dtype_code)
        if dev is not None and dev[0].dtype == want and dev[0].shape == q.shape:
            dev[0].copy_(q)
            comp = dev[1]
            if comp is not None:
                comp.zero_()
            elif dtype_code == 1:
                comp = torch.zeros_like(q)
            self._dev = (dev[0], comp)
        else:
            self._dev = (q, torch.zeros_like(q) if dtype_code == 1 else None)
        self._dev_stale = False
        self._snap = None
        self._host_current = True
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
        self._host_current = False
        self._snap = None

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

"""Solution (pyclaw/solution.py:15-420): a list of States on a Domain; frames
are written and read in the classic ascii format (fileio/ascii.py)."""
import copy

from .geometry import Dimension, Domain, Patch
from .state import State


class Solution(object):
    def __getattr__(self, key):
        if key in ("t", "num_eqn", "num_aux", "mp", "mF", "q", "p", "F", "aux", "capa",
                   "problem_data", "num_dim", "p_centers", "p_nodes", "c_centers", "c_nodes",
                   "meqn", "mbc", "grid"):
            states = self.__dict__.get("states")
            if not states:
                raise AttributeError(key)
            return getattr(states[0], key)
        raise AttributeError(key)

    def __setattr__(self, key, value):
        if key in ("t", "mp", "mF", "problem_data", "q", "aux"):
            for s in self.states:
                setattr(s, key, value)
        else:
            self.__dict__[key] = value

    state = property(lambda self: self.states[0])
    patch = property(lambda self: self.domain.patch)

    @property
    def start_frame(self):
        return self.__dict__.get("_start_frame", 0)

    def __init__(self, *arg, **kargs):
        self.__dict__["states"] = []
        self.__dict__["domain"] = None
        if len(arg) == 2:
            geom = arg[1]
            if isinstance(geom, Domain):
                self.domain = geom
            else:
                geom = list(geom) if isinstance(geom, (list, tuple)) else [geom]
                if isinstance(geom[0], Dimension):
                    self.domain = Domain(Patch(geom))
                elif isinstance(geom[0], Patch):
                    self.domain = Domain(geom)
                else:
                    raise Exception("Invalid arguments for Solution initialization.")
            if isinstance(arg[0], State):
                self.states.append(arg[0])
            elif isinstance(arg[0], (list, tuple)) and isinstance(arg[0][0], State):
                self.__dict__["states"] = list(arg[0])
            elif isinstance(arg[0], (list, tuple)) and isinstance(arg[0][0], int):
                self.states.append(State(self.domain, arg[0][0], arg[0][1]))
            elif isinstance(arg[0], int):
                self.states.append(State(self.domain, arg[0]))
            if not self.states or self.domain is None:
                raise Exception("Invalid arguments for Solution initialization.")
        elif len(arg) == 1 and isinstance(arg[0], int):
            self.read(arg[0], **kargs)       # solution.py:139-147: a frame from disk
        elif len(arg) == 0 and not kargs:
            pass
        else:
            raise Exception("Invalid arguments for Solution initialization.")

    def is_valid(self):
        return all(s.is_valid() for s in self.states)

    def __copy__(self):
        return self.__deepcopy__()

    def __deepcopy__(self, memo={}):
        result = Solution.__new__(Solution)
        result.__dict__["states"] = [copy.deepcopy(s) for s in self.states]
        result.__dict__["domain"] = copy.deepcopy(self.domain)
        return result

    def write(self, frame, path="./", file_format="ascii", file_prefix=None, write_aux=False,
              options={}, write_p=False):
        """solution.py:253-303: write the frame in each requested format (ascii)."""
        import os
        path = os.path.expandvars(os.path.expanduser(path))
        os.makedirs(path, exist_ok=True)
        formats = [file_format] if isinstance(file_format, str) else list(file_format)
        for form in formats:
            mod = _fileio(form)
            if file_prefix is None:
                mod.write(self, frame, path, write_aux=write_aux, options=options,
                          write_p=write_p)
            else:
                mod.write(self, frame, path, file_prefix=file_prefix, write_aux=write_aux,
                          options=options, write_p=write_p)

    def read(self, frame, path="./_output", file_format=None, file_prefix="fort",
             read_aux=True, options={}, **kargs):
        """solution.py:305-363: the format comes from the fort.t file when it
        names one, else from file_format; ascii or binary (binary64, binary32)."""
        import os
        from .fileio import ascii as A
        path = os.path.expandvars(os.path.expanduser(path))
        prefix = "fort" if file_prefix is None else file_prefix
        try:
            named = A.read_t(frame, path, file_prefix=prefix)[6]
        except (OSError, ValueError):
            named = None
        if named is not None:
            file_format = named
        if file_format is None:
            file_format = "ascii"
        options = dict(options)
        options["format"] = file_format       # binary / binary64 / binary32 select the type
        _fileio(file_format, reading=True).read(self, frame, path, file_prefix=prefix,
                                                read_aux=read_aux, options=options)


def _fileio(file_format, reading=False):
    if file_format == "ascii":
        from .fileio import ascii
        return ascii
    if reading and str(file_format)[:6] == "binary":
        from .fileio import binary
        return binary
    raise NotImplementedError("file format %r is not supported (ascii is; binary frames "
                              "are read)" % (file_format,))

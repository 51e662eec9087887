"""Dimension / Patch / Domain / Grid: the parts of pyclaw.geometry the stepping
path uses (geometry.py:27-820).  Cell widths and centres are computed with the
reference's formulas so that set-ups built on them give identical inputs."""
import copy

import numpy as np


class Dimension(object):
    """geometry.py:476-645: lower, upper, num_cells, delta, centers, nodes."""

    def __init__(self, lower, upper, num_cells, name="x", on_lower_boundary=None,
                 on_upper_boundary=None, units=None):
        if upper <= lower:
            raise ValueError("Upper bound must be greater than lower bound")
        self.name = name
        self.units = units
        self._lower = float(lower)
        self._upper = float(upper)
        self._num_cells = int(num_cells)
        self.on_lower_boundary = on_lower_boundary
        self.on_upper_boundary = on_upper_boundary

    lower = property(lambda self: self._lower)
    upper = property(lambda self: self._upper)
    num_cells = property(lambda self: self._num_cells)

    @property
    def delta(self):
        """geometry.py:521-524"""
        return (self.upper - self.lower) / float(self.num_cells)

    @property
    def nodes(self):
        return np.array([self.lower + i * self.delta for i in range(self.num_cells + 1)])

    @property
    def centers(self):
        """geometry.py:546-553: lower + (i+0.5)*delta, element by element."""
        return np.array([self.lower + (i + 0.5) * self.delta for i in range(self.num_cells)])

    def centers_with_ghost(self, num_ghost):
        """geometry.py:583-589: ghost centres below from lower, above from upper."""
        d = self.delta
        below = self.lower + (np.arange(-num_ghost, 0) + 0.5) * d
        above = self.upper + d * (np.arange(num_ghost) + 0.5)
        return np.concatenate((below, self.centers, above))

    def nodes_with_ghost(self, num_ghost):
        """geometry.py:591-597: ghost nodes as evenly spaced runs outside each end."""
        d = self.delta
        below = np.linspace(self.lower - num_ghost * d, self.lower - d, num_ghost)
        above = np.linspace(self.upper + d, self.upper + num_ghost * d, num_ghost)
        return np.concatenate((below, self.nodes, above))

    def __len__(self):
        return self.num_cells

    def __str__(self):
        return "Dimension %s:  (num_cells,delta,[lower,upper]) = (%s,%s,[%s,%s])" % (
            self.name, self.num_cells, self.delta, self.lower, self.upper)


def _identity_map(*xc):
    """geometry.py:13-25: the default mapping, physical = computational."""
    return list(xc)


class Grid(object):
    """geometry.py:27-470 (no gauges).  Physical coordinates are the
    computational ones mapped by grid.mapc2p (identity by default)."""

    def __init__(self, dimensions):
        if isinstance(dimensions, Dimension):
            dimensions = [dimensions]
        self._dimensions = []
        for d in dimensions:
            self.add_dimension(d)
        self.gauges = []
        self.gauge_files = []
        self.gauge_file_names = []
        self.mapc2p = _identity_map

    def add_dimension(self, dimension):
        if dimension.name in self.get_dim_attribute("name"):
            raise Exception("Unable to add dimension. A dimension of the same name: "
                            "%s, already exists." % dimension.name)
        self._dimensions.append(dimension.name)
        setattr(self, dimension.name, dimension)

    def get_dim_attribute(self, attr):
        return [getattr(getattr(self, name), attr) for name in self._dimensions]

    @property
    def num_dim(self):
        return len(self._dimensions)

    @property
    def dimensions(self):
        return [getattr(self, name) for name in self._dimensions]

    num_cells = property(lambda self: self.get_dim_attribute("num_cells"))
    lower = property(lambda self: self.get_dim_attribute("lower"))
    upper = property(lambda self: self.get_dim_attribute("upper"))
    delta = property(lambda self: self.get_dim_attribute("delta"))
    centers = property(lambda self: self.get_dim_attribute("centers"))
    nodes = property(lambda self: self.get_dim_attribute("nodes"))
    on_lower_boundary = property(lambda self: [True] * self.num_dim)
    on_upper_boundary = property(lambda self: [True] * self.num_dim)

    @property
    def c_centers(self):
        return np.meshgrid(*[d.centers for d in self.dimensions], indexing="ij")

    @property
    def c_nodes(self):
        return np.meshgrid(*[d.nodes for d in self.dimensions], indexing="ij")

    def _map(self, xc):
        f = self.mapc2p if self.mapc2p is not None else _identity_map
        return list(f(*xc))

    @property
    def p_centers(self):
        """geometry.py:303-312: mapc2p of the computational centres."""
        return self._map(self.c_centers)

    @property
    def p_nodes(self):
        """geometry.py:314-323: mapc2p of the computational nodes."""
        return self._map(self.c_nodes)

    def c_centers_with_ghost(self, num_ghost):
        return np.meshgrid(*[d.centers_with_ghost(num_ghost) for d in self.dimensions], indexing="ij")

    def p_centers_with_ghost(self, num_ghost):
        return self._map(self.c_centers_with_ghost(num_ghost))

    def c_nodes_with_ghost(self, num_ghost):
        """geometry.py:350-366"""
        return np.meshgrid(*[d.nodes_with_ghost(num_ghost) for d in self.dimensions], indexing="ij")

    def p_nodes_with_ghost(self, num_ghost):
        """geometry.py:368-369"""
        return self._map(self.c_nodes_with_ghost(num_ghost))

    def __copy__(self):
        return self.__class__(copy.deepcopy(self.dimensions))

    def __str__(self):
        return "%s-dimensional grid: %s" % (self.num_dim, ", ".join(str(d) for d in self.dimensions))


class Patch(object):
    """geometry.py:647-741."""

    def __init__(self, dimensions):
        self.level = 1
        self.patch_index = 1
        if isinstance(dimensions, Dimension):
            dimensions = [dimensions]
        # one patch covers the domain: every side is a physical boundary
        # (geometry.py:692-695)
        for dim in dimensions:
            dim.on_lower_boundary = True
            dim.on_upper_boundary = True
        self.grid = Grid(dimensions)

    def __getattr__(self, key):
        if key in ("num_dim", "dimensions", "num_cells", "lower", "upper", "delta",
                   "centers", "nodes", "p_centers", "c_centers", "p_nodes", "c_nodes"):
            return getattr(self.grid, key)
        if key in ("num_cells_global", "lower_global", "upper_global"):
            return getattr(self.grid, key[:-len("_global")])
        grid = self.__dict__.get("grid")
        if grid is not None and key in grid._dimensions:
            return getattr(grid, key)
        raise AttributeError(key)

    def __deepcopy__(self, memo={}):
        result = Patch(copy.deepcopy(self.grid.dimensions))
        result.level, result.patch_index = self.level, self.patch_index
        result.grid.mapc2p = self.grid.mapc2p
        return result


class Domain(object):
    """geometry.py:744-820: Domain(lower, upper, num_cells) or Domain(dims | patch)."""

    def __init__(self, *arg):
        if len(arg) > 1:
            lower, upper, n = arg[0], arg[1], arg[2]
            names = ["x", "y", "z"]
            dims = [Dimension(lo, up, nn, name=name) for lo, up, nn, name in zip(lower, upper, n, names)]
            self.patches = [Patch(dims)]
        else:
            geom = arg[0]
            if not isinstance(geom, (list, tuple)):
                geom = [geom]
            if isinstance(geom[0], Patch):
                self.patches = list(geom)
            elif isinstance(geom[0], Dimension):
                self.patches = [Patch(list(geom))]
            else:
                raise Exception("Invalid arguments for Domain initialization.")

    patch = property(lambda self: self.patches[0])
    grid = property(lambda self: self.patches[0].grid)
    num_dim = property(lambda self: self.patches[0].grid.num_dim)

    def __deepcopy__(self, memo={}):
        return Domain(copy.deepcopy(self.patches))

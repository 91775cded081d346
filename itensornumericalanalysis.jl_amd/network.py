"""Tree tensor networks on a digit layout (the `ITensorNetworkFunction` of the
reference, src/itensornetworkfunction.jl:18-57).

Internal layout, shared with the C-ABI of libqtt (include/qtt.h):
  * vertices are renumbered in breadth-first order from the graph's first
    vertex, so parent[v] < v and parent[0] = -1;
  * core v is a float64 array [bond to parent][site][bond to child 1]...
    (children in increasing internal id); the root's parent bond is 1;
  * operator cores are [parent][s_in][s_out][children...] with
    phi[s_out] = sum_{s_in} O[s_in, s_out] psi[s_in]
    (src/digit_inds.jl:35-58, src/elementary_operators.jl:296).
A network may live on the host (numpy cores), on the device (a libqtt handle),
or both; `cores` downloads on demand.
"""
from __future__ import annotations

import numpy as np

from .indsnetworkmap import IndsNetworkMap


class Topology:
    """Rooted, renumbered tree derived from an IndsNetworkMap."""

    def __init__(self, s: IndsNetworkMap):
        g = s.graph
        if not g.is_tree():
            raise NotImplementedError(
                "networks on graphs with loops are not supported (tree networks only)")
        root = g.vertices[0]
        order = [root]
        parent_name = {root: None}
        head = 0
        while head < len(order):
            v = order[head]
            head += 1
            for u in g.neighbors(v):
                if u not in parent_name:
                    parent_name[u] = v
                    order.append(u)
        self.names = order
        self.id = {v: i for i, v in enumerate(order)}
        nv = len(order)
        self.nv = nv
        self.parent = np.array([-1 if parent_name[v] is None else self.id[parent_name[v]]
                                for v in order], dtype=np.int32)
        self.children = [[] for _ in range(nv)]
        for v in range(1, nv):
            self.children[self.parent[v]].append(v)
        self.sdim = np.ones(nv, dtype=np.int32)
        self.vcoord = -np.ones(nv, dtype=np.int32)
        self.vdigit = np.zeros(nv, dtype=np.int32)
        for v, name in enumerate(order):
            site = s.site(name)
            if site is not None:
                self.sdim[v] = s.base
                self.vcoord[v] = site[0] - 1
                self.vdigit[v] = site[1]
        self.ncoord = s.dimension()
        self.Lc = np.array([s.ndigits(d) for d in range(1, self.ncoord + 1)], dtype=np.int32)

    def core_shape(self, v, bond):
        return (int(bond[v]), int(self.sdim[v])) + tuple(int(bond[c]) for c in self.children[v])

    def op_shape(self, v, bond):
        d = int(self.sdim[v])
        return (int(bond[v]), d, d) + tuple(int(bond[c]) for c in self.children[v])

    def is_chain(self):
        return all(len(c) <= 1 for c in self.children) and all(
            self.parent[v] == v - 1 for v in range(1, self.nv))

    def same(self, other: "Topology"):
        return (self.names == other.names and np.array_equal(self.parent, other.parent)
                and np.array_equal(self.vcoord, other.vcoord)
                and np.array_equal(self.vdigit, other.vdigit))


def bonds_of(topo: Topology, cores):
    b = np.ones(topo.nv, dtype=np.int64)
    for v in range(topo.nv):
        b[v] = cores[v].shape[0]
    return b


class ITensorNetworkFunction:
    """A function of `s.dimension()` coordinates stored as a tree tensor network."""

    def __init__(self, s: IndsNetworkMap, cores=None, topo: Topology | None = None, device=None):
        self.s = s
        self.topo = topo if topo is not None else Topology(s)
        self._cores = None
        self._device = device  # engine handle (DeviceNetwork) or None
        if cores is not None:
            self.set_cores(cores)
        elif device is None:
            raise ValueError("a network needs host cores or a device handle")

    # --- storage ---------------------------------------------------------
    def set_cores(self, cores):
        cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
        if len(cores) != self.topo.nv:
            raise ValueError("one core per vertex expected")
        bond = bonds_of(self.topo, cores)
        if bond[0] != 1:
            raise ValueError("the root core must have parent bond 1")
        for v in range(self.topo.nv):
            if cores[v].shape != self.topo.core_shape(v, bond):
                raise ValueError(f"core {v} has shape {cores[v].shape}, "
                                 f"expected {self.topo.core_shape(v, bond)}")
        self._cores = cores
        self._device = None

    @property
    def cores(self):
        if self._cores is None:
            self._cores = self._device.download()
        return self._cores

    @property
    def device(self):
        return self._device

    def attach(self, device):
        self._device = device

    def bonds(self):
        if self._cores is None:
            return self._device.bonds()
        return bonds_of(self.topo, self._cores)

    def maxlinkdim(self):
        return int(self.bonds().max())

    def copy(self):
        return ITensorNetworkFunction(self.s, [c.copy() for c in self.cores], self.topo)

    def __getitem__(self, name):
        return self.cores[self.topo.id[name]]

    def dimension(self):
        return self.s.dimension()

    def dimensions(self):
        return self.s.dimensions()

    def dimension_vertices(self, dim):
        return self.s.dimension_vertices(dim)

    # --- arithmetic (on the device, as the reference's `+` and `*`) --------
    def __add__(self, other):
        from .algebra import add
        return add(self, other)

    def __mul__(self, other):
        from .algebra import multiply
        from .builders import scale
        if isinstance(other, ITensorNetworkFunction):
            return multiply(self, other)
        return scale(self, float(other))

    __rmul__ = __mul__

    def __repr__(self):
        return f"ITensorNetworkFunction(nv={self.topo.nv}, maxlinkdim={self.maxlinkdim()})"


class TreeOperator:
    """An operator network (MPO on a chain, TTNO on a tree) on a digit layout."""

    def __init__(self, s: IndsNetworkMap, cores=None, topo: Topology | None = None, device=None):
        self.s = s
        self.topo = topo if topo is not None else Topology(s)
        self._cores = None
        self._device = device  # engine handle (DeviceOperator) or None
        if cores is not None:
            cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
            bond = bonds_of(self.topo, cores)
            for v in range(self.topo.nv):
                if cores[v].shape != self.topo.op_shape(v, bond):
                    raise ValueError(f"operator core {v} has shape {cores[v].shape}")
            self._cores = cores
        elif device is None:
            raise ValueError("an operator needs host cores or a device handle")

    @property
    def cores(self):
        if self._cores is None:
            self._cores = self._device.download()
        return self._cores

    def bonds(self):
        if self._cores is None:
            return self._device.bonds()
        return bonds_of(self.topo, self._cores)

    def maxlinkdim(self):
        return int(self.bonds().max())

    def __add__(self, other):
        from .operators import add_operators
        return add_operators(self, other)

    def __repr__(self):
        return f"TreeOperator(nv={self.topo.nv}, maxlinkdim={self.maxlinkdim()})"

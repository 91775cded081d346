"""Functions on layouts whose graph has loops (the grids of
test/test_realitensorfunction.jl:41), evaluated by belief propagation on the
device (libqtt qtt_bp_eval_host): the reference's evaluate(f, xs; alg="bp")
(src/itensornetworkfunction.jl:96-106 -> ITensorNetworks.scalar(tn; alg="bp")).

A LoopyNetworkFunction keeps one core per vertex, [site][legs...], its legs
being the vertex's edges in the order of the layout's edge list.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .indsnetworkmap import DigitMapError, IndsNetworkMap


class LoopyNetworkFunction:
    def __init__(self, s: IndsNetworkMap, cores, edge_bonds):
        self.s = s
        g = s.graph
        self.names = list(g.vertices)
        self.id = {v: i for i, v in enumerate(self.names)}
        self.edges = [(self.id[a], self.id[b]) for a, b in g.edges]
        self.edge_bonds = np.asarray(edge_bonds, dtype=np.int64)
        if len(self.edge_bonds) != len(self.edges):
            raise ValueError("one bond per edge")
        nv = len(self.names)
        self.legs = [[e for e, (a, b) in enumerate(self.edges) if v in (a, b)] for v in range(nv)]
        self.sdim = np.ones(nv, dtype=np.int32)
        self.vcoord = -np.ones(nv, dtype=np.int32)
        self.vdigit = np.zeros(nv, dtype=np.int32)
        for v, name in enumerate(self.names):
            site = s.site(name)
            if site is not None:
                self.sdim[v], self.vcoord[v], self.vdigit[v] = s.base, site[0] - 1, site[1]
        self.Lc = np.array([s.ndigits(d) for d in s.dimensions()], dtype=np.int32)
        cores = [np.ascontiguousarray(c, dtype=np.float64) for c in cores]
        for v in range(nv):
            if cores[v].shape != self.core_shape(v):
                raise ValueError(f"core {v} has shape {cores[v].shape}, expected {self.core_shape(v)}")
        self.cores = cores

    def core_shape(self, v):
        return (int(self.sdim[v]),) + tuple(int(self.edge_bonds[e]) for e in self.legs[v])

    def dimension(self):
        return self.s.dimension()

    def maxlinkdim(self):
        return int(self.edge_bonds.max()) if len(self.edge_bonds) else 1


def loopy_const(s: IndsNetworkMap, c=1.0):
    """f = c with bond 1 on every edge (the first vertex carries c)."""
    names = list(s.graph.vertices)
    cores = []
    for v, name in enumerate(names):
        d = s.base if s.site(name) is not None else 1
        deg = sum(1 for a, b in s.graph.edges if name in (a, b))
        core = np.ones((d,) + (1,) * deg) * (c if v == 0 else 1.0)
        cores.append(core)
    return LoopyNetworkFunction(s, cores, np.ones(len(s.graph.edges), dtype=np.int64))


def loopy_random(s: IndsNetworkMap, link_space: int = 2, rng=None, std=1.0):
    """Random cores, N(0, std^2) entries, bond link_space on every edge."""
    rng = rng if rng is not None else np.random.default_rng()
    bonds = np.full(len(s.graph.edges), int(link_space), dtype=np.int64)
    shapes = []
    names = list(s.graph.vertices)
    for name in names:
        d = s.base if s.site(name) is not None else 1
        deg = sum(1 for a, b in s.graph.edges if name in (a, b))
        shapes.append((d,) + (int(link_space),) * deg)
    return LoopyNetworkFunction(s, [rng.standard_normal(sh) * std for sh in shapes], bonds)


def from_tree(f) -> LoopyNetworkFunction:
    """A tree network in the loopy representation (same values): core
    [parent][site][children] becomes [site][legs in edge order]."""
    s = f.s
    t = f.topo
    g = s.graph
    names = list(g.vertices)
    edges = list(g.edges)
    bonds = f.bonds()
    eb = []
    for a, b in edges:
        ia, ib = t.id[a], t.id[b]
        eb.append(int(bonds[ia] if t.parent[ia] == ib else bonds[ib]))
    out = []
    for name in names:
        v = t.id[name]
        core = f.cores[v]
        legs = [e for e, (a, b) in enumerate(edges) if name in (a, b)]
        # axes of the tree core: parent (0), site (1), children (2..)
        axis_of = {}
        for e in legs:
            a, b = edges[e]
            other = t.id[b] if a == name else t.id[a]
            axis_of[e] = 0 if t.parent[v] == other else 2 + t.children[v].index(other)
        perm = [1] + [axis_of[e] for e in legs]
        arr = np.transpose(core, perm + ([0] if t.parent[v] < 0 else []))
        if t.parent[v] < 0:
            arr = arr[..., 0]  # the root's unit parent leg
        out.append(arr)
    return LoopyNetworkFunction(s, out, eb)


def _graph_struct(f: LoopyNetworkFunction):
    eu = np.array([a for a, _ in f.edges] or [0], dtype=np.int32)
    ev = np.array([b for _, b in f.edges] or [0], dtype=np.int32)
    keep = [eu, ev, np.ascontiguousarray(f.edge_bonds if len(f.edge_bonds) else np.ones(1, np.int64)),
            f.sdim, f.vcoord, f.vdigit, f.Lc if len(f.Lc) else np.zeros(1, np.int32)]
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    st = _capi.QttGraph(len(f.names), len(f.edges), i32(keep[0]), i32(keep[1]),
                        keep[2].ctypes.data_as(C.POINTER(C.c_int64)), i32(keep[3]), i32(keep[4]), i32(keep[5]),
                        len(f.Lc), i32(keep[6]))
    return st, keep


def evaluate_bp(f: LoopyNetworkFunction, xs, maxiter: int = 200, tol: float = 1e-13, return_sweeps=False):
    """f at points xs ([npts, ndims], or one point) by belief propagation on the device."""
    from .api import engine_context
    a = np.asarray(xs, dtype=np.float64)
    single = a.ndim <= 1
    a = np.ascontiguousarray(a.reshape(-1, f.dimension()) if a.ndim else a.reshape(1, 1))
    if a.shape[1] != f.dimension():
        raise ValueError("wrong number of coordinates")
    ctx = engine_context()
    st, keep = _graph_struct(f)
    out = np.empty(a.shape[0])
    sweeps = C.c_int(0)
    rc = _capi.lib().qtt_bp_eval_host(ctx.handle, C.byref(st), _capi._dpp(f.cores), _capi._dptr(a), a.shape[0],
                                      _capi._dptr(out), int(maxiter), float(tol), C.byref(sweeps))
    if rc == _capi.QTT_ERANGE:
        raise DigitMapError("a coordinate is negative or NaN (the digit map is defined for x >= 0)")
    ctx.check(rc)
    res = float(out[0]) if single else out
    return (res, sweeps.value) if return_sweeps else res

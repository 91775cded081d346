"""Function algebra on device-resident networks: Hadamard product, sum and
partial integration (the Fredholm route of examples/fredholm_solver.jl:43-65).

  multiply(g, f)               src/elementary_operators.jl:260-292
  add(f, g), f + g             ITensorNetworks.add (src/fixes.jl:40-95)
  partial_integrate(f, dims)   src/integration.jl:35-54, merging the integrated
                               vertices away as merge_internal_tensors
                               (src/utils.jl:99-113)
  operator_proj(f)             src/elementary_operators.jl:246-258
  map_to_zero_operator,        src/elementary_operators.jl:213-243
  map_to_zeros

Each call runs on the device (libqtt: qtt_multiply, qtt_add,
qtt_partial_integrate, qtt_operator_proj) and returns a device-resident
network or operator; the layouts of the results are planned here from the
operands' layouts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .indsnetworkmap import IndsNetworkMap
from .network import ITensorNetworkFunction, Topology, TreeOperator


def _dev(f: ITensorNetworkFunction):
    from .api import device_network
    return device_network(f)


def _wrap(ctx, s: IndsNetworkMap, topo: Topology, handle) -> ITensorNetworkFunction:
    return ITensorNetworkFunction(s, None, topo, device=_capi.DeviceNetwork(ctx, topo, handle))


def multiply(g: ITensorNetworkFunction, f: ITensorNetworkFunction, *more) -> ITensorNetworkFunction:
    """(g f)(x) = g(x) f(x).  The product lives on the union of the two layouts
    (a vertex only one of them has keeps that function's core); on common
    vertices the cores are Kronecker products, bonds multiply (g's index
    major)."""
    if more:
        return multiply(multiply(g, f), *more)
    dg, df = _dev(g), _dev(f)
    ctx = dg.ctx
    s = g.s if g.s.same_layout(f.s) else g.s.union(f.s)
    topo = g.topo if s is g.s else Topology(s)
    g_of = np.array([g.topo.id.get(n, -1) for n in topo.names], dtype=np.int32)
    f_of = np.array([f.topo.id.get(n, -1) for n in topo.names], dtype=np.int32)
    st, keep = _capi._topology_struct(topo)
    h = C.c_void_p()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    ctx.check(_capi.lib().qtt_multiply(ctx.handle, dg.handle, df.handle, C.byref(st), ip(g_of), ip(f_of),
                                       C.byref(h)))
    return _wrap(ctx, s, topo, h)


def add(f: ITensorNetworkFunction, g: ITensorNetworkFunction) -> ITensorNetworkFunction:
    """(f + g)(x): direct sum of the cores, bonds add.  Both on one layout."""
    if not f.s.same_layout(g.s):
        raise ValueError("networks on different layouts (the sum needs the same vertices)")
    df, dg = _dev(f), _dev(g)
    ctx = df.ctx
    h = C.c_void_p()
    ctx.check(_capi.lib().qtt_add(ctx.handle, df.handle, dg.handle, C.byref(h)))
    return _wrap(ctx, f.s, f.topo, h)


def partial_integrate(f: ITensorNetworkFunction, dims, merge_vertices: bool = True,
                      take_sum: bool = False) -> ITensorNetworkFunction:
    """The integral of f over the coordinates `dims` (left Riemann sum, or the
    plain sum with take_sum): a function of the remaining coordinates.  With
    merge_vertices the integrated vertices (and vertices without digits) are
    contracted into the remaining ones, whose layout is then the reduced
    layout of the remaining dimensions; without, the network keeps its
    vertices and the integrated ones carry no digit."""
    dims = [dims] if isinstance(dims, int) else sorted(set(int(d) for d in dims))
    s = f.s
    for d in dims:
        if not 1 <= d <= s.dimension():
            raise ValueError(f"no dimension {d}")
    rest = [d for d in s.dimensions() if d not in dims and s.ndigits(d) > 0]
    if merge_vertices:
        if not rest:
            raise ValueError("every dimension is integrated: use integrate(f)")
        s_new = s.reduced(rest)
    else:
        s_new = s.without_dimensions(dims)
    topo = Topology(s_new)
    out_vertex = np.array([f.topo.id[n] for n in topo.names], dtype=np.int32)
    mask = np.zeros(max(f.topo.ncoord, 1), dtype=np.int32)
    for d in dims:
        mask[d - 1] = 1
    df = _dev(f)
    ctx = df.ctx
    st, keep = _capi._topology_struct(topo)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    h = C.c_void_p()
    w = 1.0 if take_sum else 1.0 / s.base
    ctx.check(_capi.lib().qtt_partial_integrate(ctx.handle, df.handle, ip(mask), w, C.byref(st), ip(out_vertex),
                                                C.byref(h)))
    return _wrap(ctx, s_new, topo, h)


def operator_proj(f: ITensorNetworkFunction) -> TreeOperator:
    """|f><delta|: the diagonal operator with operate(O, g) = f g, built on the
    device from f's cores (src/elementary_operators.jl:246-258)."""
    df = _dev(f)
    ctx = df.ctx
    h = C.c_void_p()
    ctx.check(_capi.lib().qtt_operator_proj(ctx.handle, df.handle, C.byref(h)))
    return TreeOperator(f.s, None, f.topo, device=_capi.DeviceOperator(ctx, f.topo, handle=h))


def map_to_zero_operator(s: IndsNetworkMap, xs, dims=None) -> TreeOperator:
    """The operator that sets a function to 0 at the grid points xs (one
    coordinate each, in dimensions dims, default 1) and leaves it unchanged
    elsewhere: the projection of 1 - sum of the points' indicators, with
    crossings of points in different dimensions added back
    (src/elementary_operators.jl:213-229)."""
    from .builders import delta_kernel
    xs = [xs] if np.isscalar(xs) else list(xs)
    dims = [1] * len(xs) if dims is None else ([dims] if np.isscalar(dims) else list(dims))
    if len(dims) != len(xs):
        raise ValueError("one dimension per point")
    kernel = delta_kernel(s, [[x] for x in xs], [[d] for d in dims], remove_overlap=True, coeff=-1.0,
                          include_identity=True)
    return operator_proj(kernel)


def map_to_zeros(f: ITensorNetworkFunction, xs, dims=None, cutoff=None, maxdim=None) -> ITensorNetworkFunction:
    """f with the grid points xs (dimensions dims) mapped to 0
    (src/elementary_operators.jl:231-243)."""
    from .api import operate
    return operate(map_to_zero_operator(f.s, xs, dims), f, cutoff=cutoff, maxdim=maxdim)

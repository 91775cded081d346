"""Host-side construction of input networks (built once, then uploaded).

Restates the reference builders by their function values on the grid:
  const_itn  src/elementary_functions.jl:16-26   f = c
  exp_itn    src/elementary_functions.jl:30-53   f = c exp(k x + a), bond 1
  cos_itn    src/elementary_functions.jl:109-119 f = c cos(k x + a), bond 2
  sin_itn    src/elementary_functions.jl:125-135 f = c sin(k x + a), bond 2
  cosh/sinh  src/elementary_functions.jl:57-83   sums of two exponentials
  random_itn src/elementary_functions.jl:209-221 (values from numpy's PCG64
             instead of Julia's Xoshiro stream)
  poly_itn   src/elementary_functions.jl:139-206 f = c sum_i coeffs[i] (k x)^i,
             bond len(coeffs), built as binomial transfer cores (below)
  delta_p    src/elementary_functions.jl:224-262 indicator of grid points
  delta_kernel src/elementary_functions.jl:264-324 1 - sum of point indicators,
             overlaps of points in different dimensions added back
add_host / multiply_host assemble synthetic inputs on the host before upload;
sums and products of networks otherwise run on the device (algebra.py).
x is the grid value sum_j digit_j 2^-j of the chosen dimension.  The reference
builds cos/sin as complex sums of two exponentials; here the same values come
from real rank-2 cores that compose the complex phase exp(i k x) as a 2-vector
(real, imaginary) -- a real gauge of the same network.
"""
from __future__ import annotations

import math

import numpy as np

from .indsnetworkmap import IndsNetworkMap
from .network import ITensorNetworkFunction, Topology


def _site_values(topo: Topology, v: int, dim: int):
    """Grid contribution s * 2^-j of vertex v's digit if it belongs to `dim`, else None."""
    if topo.vcoord[v] == dim - 1:
        j = int(topo.vdigit[v])
        return np.array([s * 2.0 ** (-j) for s in range(int(topo.sdim[v]))])
    return None


def _product_state(s: IndsNetworkMap, site_factor, root_scale=1.0):
    """Bond-1 network with core v = site_factor(v) (a vector over the site)."""
    topo = Topology(s)
    cores = []
    for v in range(topo.nv):
        vals = np.asarray(site_factor(topo, v), dtype=np.float64)
        if v == 0:
            vals = vals * root_scale
        cores.append(vals.reshape(topo.core_shape(v, np.ones(topo.nv, dtype=np.int64))))
    return ITensorNetworkFunction(s, cores, topo)


def const_itn(s: IndsNetworkMap, c=1.0, linkdim: int = 1):
    """f = c.  With linkdim > 1 every core is a delta over its bonds (the
    reference's c_tensor, src/utils.jl:42-47) and the root carries c / linkdim.
    On a layout with loops: bond 1, evaluated by belief propagation."""
    if not s.graph.is_tree():
        from .loopy import loopy_const
        if linkdim != 1:
            raise NotImplementedError("a constant with linkdim > 1 on a graph with loops needs exact "
                                      "contraction, which the engine does not do (only belief propagation)")
        return loopy_const(s, c)
    topo = Topology(s)
    chi = int(linkdim)
    bond = np.full(topo.nv, chi, dtype=np.int64)
    bond[0] = 1
    cores = []
    for v in range(topo.nv):
        shape = topo.core_shape(v, bond)
        core = np.zeros(shape)
        m = len(topo.children[v])
        if v == 0 and m == 0:
            core[...] = c
        else:
            for k in range(chi):
                p = 0 if v == 0 else k
                core[(p, slice(None)) + (k,) * m] = 1.0
            if v == 0:
                core *= c / chi
        cores.append(core)
    return ITensorNetworkFunction(s, cores, topo)


def exp_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    """f = c exp(k x_dim + a) as a product state (bond 1)."""
    if isinstance(k, complex) or isinstance(a, complex) or isinstance(c, complex):
        raise TypeError("complex exponentials: use cos_itn/sin_itn (real networks only)")
    L = s.ndigits(dim)

    def fac(topo, v):
        xv = _site_values(topo, v, dim)
        if xv is None:
            return np.ones(int(topo.sdim[v]))
        return np.exp(k * xv) * np.exp(a / L)

    return _product_state(s, fac, root_scale=c)


def _phase_itn(s: IndsNetworkMap, k, a, c, dim, part):
    """Real rank-2 network of part(c exp(i (k x + a))), part in {'re', 'im'}."""
    topo = Topology(s)
    bond = np.full(topo.nv, 2, dtype=np.int64)
    bond[0] = 1
    C = c * np.exp(1j * a)
    cores = []
    for v in range(topo.nv):
        xv = _site_values(topo, v, dim)
        d = int(topo.sdim[v])
        w = np.ones(d, dtype=complex) if xv is None else np.exp(1j * k * xv)
        m = len(topo.children[v])
        shape = topo.core_shape(v, bond)
        core = np.zeros(shape)
        for s_ in range(d):
            for idx in np.ndindex(*((2,) * m)):
                z = w[s_] * (1j ** sum(idx))
                if v == 0:
                    z = C * z
                    core[(0, s_) + idx] = z.imag if part == "im" else z.real
                else:
                    core[(0, s_) + idx] = z.real
                    core[(1, s_) + idx] = z.imag
        cores.append(core)
    return ITensorNetworkFunction(s, cores, topo)


def cos_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    """f = c cos(k x_dim + a), bond 2."""
    return _phase_itn(s, k, a, c, dim, "re")


def sin_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    """f = c sin(k x_dim + a), bond 2."""
    return _phase_itn(s, k, a, c, dim, "im")


def cosh_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    return add_host(exp_itn(s, k=k, a=a, c=0.5 * c, dim=dim), exp_itn(s, k=-k, a=-a, c=0.5 * c, dim=dim))


def sinh_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    return add_host(exp_itn(s, k=k, a=a, c=0.5 * c, dim=dim), exp_itn(s, k=-k, a=-a, c=-0.5 * c, dim=dim))


def random_itn(s: IndsNetworkMap, link_space: int = 1, seed=None, rng=None, std=1.0,
               normalize_cores=False):
    """Dense random cores, entries N(0, std^2), every bond = link_space.  Cores are
    drawn in internal vertex order from `numpy.random.default_rng(seed)`.  On a
    layout with loops: a LoopyNetworkFunction (cores in graph vertex order)."""
    if not s.graph.is_tree():
        from .loopy import loopy_random
        return loopy_random(s, link_space, rng if rng is not None else np.random.default_rng(seed), std)
    topo = Topology(s)
    if rng is None:
        rng = np.random.default_rng(seed)
    bond = np.full(topo.nv, int(link_space), dtype=np.int64)
    bond[0] = 1
    cores = []
    for v in range(topo.nv):
        core = rng.standard_normal(topo.core_shape(v, bond)) * std
        if normalize_cores:
            core /= np.linalg.norm(core)
        cores.append(core)
    return ITensorNetworkFunction(s, cores, topo)


def scale(f: ITensorNetworkFunction, c: float):
    cores = [x.copy() for x in f.cores]
    cores[0] *= c
    return ITensorNetworkFunction(f.s, cores, f.topo)



def _check_same(f, g):
    if not f.topo.same(g.topo):
        raise ValueError("networks live on different digit layouts")


def add_host(f: ITensorNetworkFunction, g: ITensorNetworkFunction):
    """Direct sum f + g of two host networks (every bond dimension added), for
    assembling synthetic inputs before upload; the product path is algebra.add."""
    _check_same(f, g)
    topo = f.topo
    bf, bg = f.bonds(), g.bonds()
    bond = bf + bg
    bond[0] = 1
    cores = []
    for v in range(topo.nv):
        A, B = f.cores[v], g.cores[v]
        C = np.zeros(topo.core_shape(v, bond))
        m = len(topo.children[v])
        if v == 0:
            if m == 0:
                C[...] = A + B
            else:
                C[(slice(0, 1), slice(None)) + tuple(slice(0, bf[c]) for c in topo.children[v])] = A
                C[(slice(0, 1), slice(None)) + tuple(slice(bf[c], bond[c]) for c in topo.children[v])] = B
        else:
            C[(slice(0, bf[v]), slice(None)) + tuple(slice(0, bf[c]) for c in topo.children[v])] = A
            C[(slice(bf[v], bond[v]), slice(None)) + tuple(slice(bf[c], bond[c]) for c in topo.children[v])] = B
        cores.append(C)
    return ITensorNetworkFunction(f.s, cores, topo)


def multiply_host(f: ITensorNetworkFunction, g: ITensorNetworkFunction):
    """Hadamard product (f g)(x) = f(x) g(x) of two host networks (bonds
    multiply, merged index (i_f, i_g)), for assembling synthetic inputs before
    upload; the product path is algebra.multiply."""
    _check_same(f, g)
    topo = f.topo
    bond = f.bonds() * g.bonds()
    cores = []
    for v in range(topo.nv):
        A, B = f.cores[v], g.cores[v]
        m = len(topo.children[v])
        # A[p, s, c...] B[q, s, e...] -> C[(p,q), s, (c,e)...]
        la = "abcdefgh"[: m]
        lb = "ijklmnop"[: m]
        spec_a = "P" + "s" + la
        spec_b = "Q" + "s" + lb
        out = "PQs" + "".join(x + y for x, y in zip(la, lb))
        C = np.einsum(f"{spec_a},{spec_b}->{out}", A, B)
        cores.append(C.reshape(topo.core_shape(v, bond)))
    return ITensorNetworkFunction(f.s, cores, topo)


def poly_itn(s: IndsNetworkMap, coeffs, dim: int = 1, k=1.0, c=1.0):
    """f(x) = c * sum_i coeffs[i] (k x)^i along dimension `dim`, on any tree.

    Edge (v, parent) carries the power q = 0..n-1 of S_v, the grid value
    contributed by the digits of `dim` in v's subtree; since S_v = t_v + sum
    over children of S_c, core v holds the multinomial expansion
        core[q][s][q_1]...[q_m] = q! / (q_0! q_1! ... q_m!) t_v(s)^q_0,
    q_0 = q - sum q_i >= 0, with t_v(s) = s 2^-j for digit j of `dim` (else 0,
    so only q_0 = 0 contributes).  The root sums the powers with the
    coefficients.  Exact, bond len(coeffs) on every edge."""
    coeffs = [float(a) * float(k) ** i for i, a in enumerate(coeffs)]
    n = len(coeffs)
    if n == 1:
        return const_itn(s, c=c * coeffs[0])
    topo = Topology(s)
    bond = np.full(topo.nv, n, dtype=np.int64)
    bond[0] = 1
    fact = [math.factorial(i) for i in range(n)]
    cores = []
    for v in range(topo.nv):
        ch = topo.children[v]
        t = _site_values(topo, v, dim)
        sd = int(topo.sdim[v])
        tv = t if t is not None else np.zeros(sd)
        core = np.zeros(topo.core_shape(v, bond))
        for qs in np.ndindex(*([n] * len(ch))):
            qsum = sum(qs)
            for q in range(qsum, n):
                q0 = q - qsum
                coef = fact[q] / (fact[q0] * np.prod([fact[x] for x in qs]))
                vals = coef * tv ** q0  # 0^0 = 1
                if v == 0:
                    core[(0, slice(None)) + tuple(qs)] += coeffs[q] * c * vals
                else:
                    core[(q, slice(None)) + tuple(qs)] = vals
        cores.append(core)
    return ITensorNetworkFunction(s, cores, topo)


def _delta_point(s: IndsNetworkMap, xs, dims, topo: Topology):
    """Bond-1 indicator of one point: one-hot digits of the given dimensions,
    constant 1 over every other digit (src/elementary_functions.jl:224-241)."""
    vals = s.calculate_ind_values(list(xs), list(dims))
    cores = []
    ones = np.ones(topo.nv, dtype=np.int64)
    for v in range(topo.nv):
        name = topo.names[v]
        d = int(topo.sdim[v])
        vec = np.zeros(d)
        if name in vals:
            vec[vals[name]] = 1.0
        else:
            vec[:] = 1.0
        cores.append(vec.reshape(topo.core_shape(v, ones)))
    return ITensorNetworkFunction(s, cores, topo)


def delta_p(s: IndsNetworkMap, points, dims=None):
    """Indicator of grid points (delta_p).  `points` is one point (a number or
    a list of coordinates, with `dims` the dimensions they give, default
    1..len) or a list of points (with a list of dimension lists); several
    points are summed (bond = number of points)."""
    topo = Topology(s)
    if np.isscalar(points):
        points = [points]
    if len(points) and not np.isscalar(points[0]):
        dims = dims if dims is not None else [list(range(1, len(p) + 1)) for p in points]
        if len(points) != len(dims) or not points:
            raise ValueError("one dimension list per point")
        f = None
        for p, d in zip(points, dims):
            t = _delta_point(s, p, d, topo)
            f = t if f is None else add_host(f, t)
        return f
    dims = dims if dims is not None else list(range(1, len(points) + 1))
    return _delta_point(s, points, dims, topo)


def _overlaps(points, dims):
    """Points where two of the given ones meet (src/elementary_functions.jl:281-317):
    two points of the same dimensions meet only if equal; points of different
    dimensions meet at the point with both sets of coordinates, unless they
    disagree on a shared dimension.  The shared-dimension comparison takes
    the coordinates in their original order, as the reference does."""
    out_p, out_d = [], []
    for i in range(len(points)):
        p1, d1 = list(points[i]), list(dims[i])
        for j in range(i + 1, len(points)):
            p2, d2 = list(points[j]), list(dims[j])
            if len(d1) == len(d2) and all(np.isclose(a, b) for a, b in zip(d1, d2)) and not (
                    len(p1) == len(p2) and all(np.isclose(a, b) for a, b in zip(p1, p2))):
                continue
            ps_, ds_ = p1 + p2, d1 + d2
            order = sorted(range(len(ds_)), key=lambda k: ds_[k])  # stable, as sortperm
            ps, ds = [ps_[order[0]]], [ds_[order[0]]]
            for k in range(1, len(ds_)):
                if ds_[order[k]] != ds_[order[k - 1]]:
                    ps.append(ps_[order[k]])
                    ds.append(ds_[order[k]])
                    continue
                if np.isclose(ps_[k], ps_[k - 1]):
                    continue
                ps, ds = [], []
                break
            if ds:
                out_p.append(ps)
                out_d.append(ds)
    return out_p, out_d


def delta_kernel(s: IndsNetworkMap, points, dims=None, remove_overlap: bool = True, coeff: float = -1.0,
                 include_identity: bool = True):
    """coeff * (sum of point indicators) [+ 1] [- coeff * indicators of the
    points where two of them meet] (src/elementary_functions.jl:264-324)."""
    dims = dims if dims is not None else [list(range(1, len(p) + 1)) for p in points]
    psi = scale(delta_p(s, points, dims), coeff)
    if include_identity:
        psi = add_host(const_itn(s), psi)
    if remove_overlap and len(points) > 1:
        op, od = _overlaps(points, dims)
        if op:
            psi = add_host(psi, scale(delta_p(s, op, od), -coeff))
    return psi

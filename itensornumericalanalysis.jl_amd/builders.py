"""Host-side construction of input networks (built once, then uploaded).

Restates the reference builders by their function values on the grid:
  const_itn  src/elementary_functions.jl:16-26   f = c
  exp_itn    src/elementary_functions.jl:30-53   f = c exp(k x + a), bond 1
  cos_itn    src/elementary_functions.jl:109-119 f = c cos(k x + a), bond 2
  sin_itn    src/elementary_functions.jl:125-135 f = c sin(k x + a), bond 2
  cosh/sinh  src/elementary_functions.jl:57-83   sums of two exponentials
  random_itn src/elementary_functions.jl:209-221 (values from numpy's PCG64
             instead of Julia's Xoshiro stream)
  add        ITensorNetworks `+` (direct sum on every bond)
  multiply   src/elementary_operators.jl:260-292 (per-vertex Hadamard product)
x is the grid value sum_j digit_j 2^-j of the chosen dimension.  The reference
builds cos/sin as complex sums of two exponentials; here the same values come
from real rank-2 cores that compose the complex phase exp(i k x) as a 2-vector
(real, imaginary) -- a real gauge of the same network.
"""
from __future__ import annotations

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
    reference's c_tensor, src/utils.jl:42-47) and the root carries c / linkdim."""
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
    return exp_itn(s, k=k, a=a, c=0.5 * c, dim=dim) + exp_itn(s, k=-k, a=-a, c=0.5 * c, dim=dim)


def sinh_itn(s: IndsNetworkMap, k=1.0, a=0.0, c=1.0, dim: int = 1):
    return exp_itn(s, k=k, a=a, c=0.5 * c, dim=dim) + exp_itn(s, k=-k, a=-a, c=-0.5 * c, dim=dim)


def random_itn(s: IndsNetworkMap, link_space: int = 1, seed=None, rng=None, std=1.0,
               normalize_cores=False):
    """Dense random cores, entries N(0, std^2), every bond = link_space.  Cores are
    drawn in internal vertex order from `numpy.random.default_rng(seed)`."""
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


def add(f: ITensorNetworkFunction, g: ITensorNetworkFunction):
    """Direct sum: f + g with every bond dimension added."""
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


def multiply(f: ITensorNetworkFunction, g: ITensorNetworkFunction):
    """Hadamard product (f g)(x) = f(x) g(x): bonds multiply, merged index (i_f, i_g)."""
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

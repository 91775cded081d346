"""Operator networks (host-side construction, built once per operator).

Restates src/elementary_operators.jl:32-194 and the "Digit" site ops of
src/digit_inds.jl:35-58.  An operator is first written as the same sum of
operator strings as the reference's OpSum (forward/backward carry strings,
boundary terms, identity), then compiled to a tree operator network by a
direct sum of the rank-1 strings followed by an SVD compression that only
drops round-off (the role of `ttn(::OpSum)` in ITensorNetworks).  The linear
map is exactly the reference's; the bond gauge is this compiler's own.

Op convention: O[s_in, s_out] with phi[s_out] = sum_{s_in} O[s_in, s_out] psi[s_in]
  "D+"  O[1,0] = 1      "D-"  O[0,1] = 1
  "Ddn" O[0,0] = 1      "Dup" O[1,1] = 1      "I" identity
"""
from __future__ import annotations

import numpy as np

from .indsnetworkmap import IndsNetworkMap
from .network import Topology, TreeOperator

DEFAULT_BOUNDARY = "Dirichlet"

_OPS = {
    "D+": np.array([[0.0, 0.0], [1.0, 0.0]]),
    "D-": np.array([[0.0, 1.0], [0.0, 0.0]]),
    "Ddn": np.array([[1.0, 0.0], [0.0, 0.0]]),
    "Dup": np.array([[0.0, 0.0], [0.0, 1.0]]),
    "I": np.eye(2),
}


# ---------------------------------------------------------------------------
# operator strings (the OpSum of the reference)
# ---------------------------------------------------------------------------
def boundary_terms(s: IndsNetworkMap, boundary: str, dim: int, isfwd: bool, n: int = 0):
    """src/elementary_operators.jl:32-57."""
    L = s.ndigits(dim)
    if boundary == "Neumann":
        op = "Dup" if isfwd else "Ddn"
    elif boundary == "Periodic":
        op = "D-" if isfwd else "D+"
    elif boundary == "Dirichlet":
        return []
    else:
        raise ValueError(f"unknown boundary {boundary!r}")
    return [(1.0, {s.vertex(dim, j): op for j in range(1, L - n + 1)})]


def point_to_opsum(s: IndsNetworkMap, x: float, dim: int = 1):
    """The operator string projecting dimension `dim` onto the digits of x:
    "Dup" where x's digit is 1, "Ddn" where it is 0
    (src/elementary_operators.jl:200-210).  compile_terms makes it an operator."""
    vals = s.calculate_ind_values([x], [dim])
    return [(1.0, {v: ("Dup" if vals[v] == 1 else "Ddn") for v in s.dimension_vertices(dim)})]


def _shift_terms(s, dim, boundary, n, first, carry, isfwd):
    L = s.ndigits(dim)
    if L - n < 1:
        raise ValueError("shift larger than the grid")
    terms = [(1.0, {s.vertex(dim, L - n): first})]
    string = {}
    for i in range(L - n, 1, -1):
        string = {k: v for k, v in string.items()}
        string[s.vertex(dim, i)] = carry
        t = dict(string)
        t[s.vertex(dim, i - 1)] = first
        terms.append((1.0, t))
    return terms + boundary_terms(s, boundary, dim, isfwd, n)


def forward_shift_terms(s, dim=1, boundary=DEFAULT_BOUNDARY, n=0):
    """phi(x) = psi(x + (n+1) h) (src/elementary_operators.jl:59-80)."""
    return _shift_terms(s, dim, boundary, n, "D+", "D-", True)


def backward_shift_terms(s, dim=1, boundary=DEFAULT_BOUNDARY, n=0):
    """phi(x) = psi(x - (n+1) h) (src/elementary_operators.jl:82-103)."""
    return _shift_terms(s, dim, boundary, n, "D-", "D+", False)


def stencil_terms(s, shifts, dim=1, left_boundary=DEFAULT_BOUNDARY,
                  right_boundary=DEFAULT_BOUNDARY):
    """shifts = [x+2h, x+h, x, x-h, x-2h] coefficients (src/elementary_operators.jl:122-150)."""
    if len(shifts) != 5:
        raise ValueError("a stencil has 5 coefficients")
    terms = [(float(shifts[2]), {})]
    for i, n in ((0, 1), (1, 0)):
        if shifts[i] != 0:
            terms += [(shifts[i] * c, t) for c, t in forward_shift_terms(s, dim, right_boundary, n)]
    for i, n in ((3, 0), (4, 1)):
        if shifts[i] != 0:
            terms += [(shifts[i] * c, t) for c, t in backward_shift_terms(s, dim, left_boundary, n)]
    return terms


# ---------------------------------------------------------------------------
# compilation to a tree operator network
# ---------------------------------------------------------------------------
def _terms_to_cores(topo: Topology, terms):
    T = len(terms)
    bond = np.full(topo.nv, T, dtype=np.int64)
    bond[0] = 1
    cores = []
    for v in range(topo.nv):
        name = topo.names[v]
        d = int(topo.sdim[v])
        m = len(topo.children[v])
        core = np.zeros(topo.op_shape(v, bond))
        for t, (coef, ops) in enumerate(terms):
            o = _OPS[ops[name]] if name in ops else np.eye(d)
            if v == 0:
                o = coef * o
                idx = (0, slice(None), slice(None)) + (t,) * m
            else:
                idx = (t, slice(None), slice(None)) + (t,) * m
            if v == 0 and m == 0:
                core[0] += o
            else:
                core[idx] = o
        cores.append(core)
    return cores


def _axis(topo, x, u):
    if topo.parent[x] == u:
        return 0
    return 3 + topo.children[x].index(u)


def _move(topo, cores, x, u, svd, rel_cut):
    """Factor core x along its bond to u (QR, or SVD dropping round-off), keep the
    isometry in x and push the remainder into u."""
    ax, au = _axis(topo, x, u), _axis(topo, u, x)
    X = np.moveaxis(cores[x], ax, -1)
    shp = X.shape
    M = X.reshape(-1, shp[-1])
    if svd:
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
        p = S ** 2
        tot = p.sum()
        keep = max(1, int(np.sum(p > rel_cut * tot))) if tot > 0 else 1
        Q, R = U[:, :keep], S[:keep, None] * Vt[:keep]
    else:
        Q, R = np.linalg.qr(M)
    cores[x] = np.moveaxis(Q.reshape(shp[:-1] + (Q.shape[1],)), -1, ax)
    cores[u] = np.moveaxis(np.tensordot(R, cores[u], axes=([1], [au])), 0, au)


def _compress(topo: Topology, cores, rel_cut=1e-24):
    """Orthogonalise toward the root, then SVD every edge from the root down."""
    cores = [c.copy() for c in cores]
    if topo.nv == 1:
        return cores
    for v in range(topo.nv - 1, 0, -1):
        _move(topo, cores, v, int(topo.parent[v]), False, 0.0)

    def down(v):
        for c in topo.children[v]:
            _move(topo, cores, v, c, True, rel_cut)
            down(c)
            _move(topo, cores, c, v, False, 0.0)

    down(0)
    return cores


def compile_terms(s: IndsNetworkMap, terms, topo: Topology | None = None) -> TreeOperator:
    topo = topo if topo is not None else Topology(s)
    if not terms:
        terms = [(0.0, {})]
    cores = _compress(topo, _terms_to_cores(topo, terms))
    return TreeOperator(s, cores, topo)


def _scale_dim(op: TreeOperator, dim: int, factor: float):
    cores = [c.copy() for c in op.cores]
    for name in op.s.dimension_vertices(dim):
        cores[op.topo.id[name]] *= factor
    return TreeOperator(op.s, cores, op.topo)


def add_operators(A: TreeOperator, B: TreeOperator) -> TreeOperator:
    """Direct sum of two operator networks followed by round-off compression."""
    if not A.topo.same(B.topo):
        raise ValueError("operators live on different digit layouts")
    topo = A.topo
    ba, bb = A.bonds(), B.bonds()
    bond = ba + bb
    bond[0] = 1
    cores = []
    for v in range(topo.nv):
        C = np.zeros(topo.op_shape(v, bond))
        ch = topo.children[v]
        sa = tuple(slice(0, ba[c]) for c in ch)
        sb = tuple(slice(ba[c], bond[c]) for c in ch)
        if v == 0:
            if not ch:
                C[...] = A.cores[0] + B.cores[0]
            else:
                C[(slice(0, 1), slice(None), slice(None)) + sa] = A.cores[0]
                C[(slice(0, 1), slice(None), slice(None)) + sb] = B.cores[0]
        else:
            C[(slice(0, ba[v]), slice(None), slice(None)) + sa] = A.cores[v]
            C[(slice(ba[v], bond[v]), slice(None), slice(None)) + sb] = B.cores[v]
        cores.append(C)
    return TreeOperator(A.s, _compress(topo, cores), topo)


# ---------------------------------------------------------------------------
# public builders (names as in the reference)
# ---------------------------------------------------------------------------
def forward_shift_op(s, dim=1, boundary=DEFAULT_BOUNDARY, n=0):
    return compile_terms(s, forward_shift_terms(s, dim, boundary, n))


def backward_shift_op(s, dim=1, boundary=DEFAULT_BOUNDARY, n=0):
    return compile_terms(s, backward_shift_terms(s, dim, boundary, n))


def stencil(s, shifts, delta_power, dim=1, left_boundary=DEFAULT_BOUNDARY,
            right_boundary=DEFAULT_BOUNDARY, scale=True):
    """src/elementary_operators.jl:122-160: the stencil times b^delta_power on
    every vertex of `dim` (an overall h^-delta_power)."""
    op = compile_terms(s, stencil_terms(s, shifts, dim, left_boundary, right_boundary))
    if scale:
        op = _scale_dim(op, dim, float(s.base) ** delta_power)
    return op


def first_derivative_operator(s, **kw):
    return stencil(s, [0.0, 0.5, 0.0, -0.5, 0.0], 1, **kw)


def second_derivative_operator(s, **kw):
    return stencil(s, [0.0, 1.0, -2.0, 1.0, 0.0], 2, **kw)


def third_derivative_operator(s, **kw):
    return stencil(s, [0.5, -1.0, 0.0, 1.0, -0.5], 3, **kw)


def fourth_derivative_operator(s, **kw):
    return stencil(s, [1.0, -4.0, 6.0, -4.0, 1.0], 4, **kw)


def laplacian_operator(s, dims=None, boundary=None, **kw):
    """Sum of second derivatives over `dims` (src/elementary_operators.jl:178-189)."""
    if boundary is not None:
        kw = dict(kw, left_boundary=boundary, right_boundary=boundary)
    dims = list(dims) if dims is not None else s.dimensions()
    op = second_derivative_operator(s, dim=dims[0], **kw)
    for d in dims[1:]:
        op = op + second_derivative_operator(s, dim=d, **kw)
    return op


def identity_operator(s):
    return compile_terms(s, [(1.0, {})])


def fredholm_operator(s: IndsNetworkMap, dim: int = 1, kernel_rate: float = 1.0,
                      weight: bool = True) -> TreeOperator:
    """K f (x) = 2^-L sum_y exp(-rate |x - y|) f(y) on dimension `dim`: the
    discretised Fredholm kernel operator exp(-|x-y|) (the integral-operator
    pattern of examples/fredholm_solver.jl:43-65).

    Exact automaton of bond 3 built from the most significant digit down:
    state 0 = digits equal so far, 1 = x > y decided, 2 = x < y decided; once
    decided each later digit contributes exp(-+rate (x_j - y_j) 2^-j).  Requires a
    chain whose vertex order visits the digits of `dim` from most to least
    significant (other vertices pass the state through)."""
    topo = Topology(s)
    if not topo.is_chain():
        raise NotImplementedError("the kernel operator is built on chains")
    digs = [(v, int(topo.vdigit[v])) for v in range(topo.nv) if topo.vcoord[v] == dim - 1]
    if [j for _, j in digs] != list(range(1, len(digs) + 1)):
        raise NotImplementedError("digits of the dimension must run MSB first along the chain")
    nv = topo.nv
    w = 0.5 if weight else 1.0
    first = digs[0][0]
    last = digs[-1][0]
    bond = np.ones(nv, dtype=np.int64)
    # edge (v, v-1) carries 3 states between the first and last digit vertex
    for v in range(1, nv):
        if first < v <= last:
            bond[v] = 3
    cores = []
    for v in range(nv):
        d = int(topo.sdim[v])
        shape = topo.op_shape(v, bond)
        core = np.zeros(shape)
        left, right = bond[v], (bond[v + 1] if v + 1 < nv else 1)
        if topo.vcoord[v] != dim - 1:
            # pass-through: identity on the site, identity on the state
            for st in range(max(left, right) if left == right else 1):
                idx = (st if left > 1 else 0, slice(None), slice(None))
                idx = idx + ((st if right > 1 else 0,) if v + 1 < nv else ())
                core[idx] = np.eye(d)
            cores.append(core)
            continue
        j = int(topo.vdigit[v])
        hj = 2.0 ** (-j)
        for st_in in range(left):
            for y in range(2):
                for x in range(2):
                    if left == 1 or st_in == 0:
                        if x == y:
                            st_out, val = 0, 1.0
                        elif x > y:
                            st_out, val = 1, np.exp(-kernel_rate * hj)
                        else:
                            st_out, val = 2, np.exp(-kernel_rate * hj)
                    elif st_in == 1:
                        st_out, val = 1, np.exp(-kernel_rate * (x - y) * hj)
                    else:
                        st_out, val = 2, np.exp(kernel_rate * (x - y) * hj)
                    val *= w
                    if v == last:
                        idx = (st_in, y, x) + ((0,) if v + 1 < nv else ())
                        core[idx] += val
                    else:
                        core[(st_in, y, x, st_out)] += val
        cores.append(core)
    return TreeOperator(s, cores, topo)

"""ctypes wrapper of the C oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Every function takes the raw network description (parent, sdim, cores in the
[parent][site][children...] layout) so it never depends on the engine's classes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        lp = C.POINTER(C.c_long)
        dpp = C.POINTER(dp)
        L.or_digit.argtypes = [C.c_double, C.c_int, C.c_int]
        L.or_eval.argtypes = [C.c_int, ip, ip, lp, dpp, ip, ip, C.c_int, ip, dp, C.c_long, dp, C.c_int]
        L.or_integrate.argtypes = [C.c_int, ip, ip, lp, dpp, C.c_double, dp]
        L.or_inner.argtypes = [C.c_int, ip, ip, lp, dpp, lp, dpp, dp]
        L.or_contract.argtypes = [C.c_int, ip, ip, lp, dpp, lp, dpp, dpp]
        L.or_contract_adjoint.argtypes = [C.c_int, ip, ip, lp, dpp, lp, dpp, dpp, dpp, dpp]
        L.or_truncate.argtypes = [C.c_int, ip, ip, lp, dpp, C.c_double, C.c_long, dp]
        L.or_eval_grad.argtypes = [C.c_int, ip, ip, lp, dpp, ip, ip, C.c_int, ip, dp, C.c_long, dp, dpp,
                                   C.c_int]
        L.or_svd.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double]
        L.or_qr.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_trunc_rule.argtypes = [dp, C.c_long, C.c_double, C.c_long, dp]
        L.or_trunc_rule.restype = C.c_long
        _LIB = L
    return _LIB


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _l(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_long))


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _dpp(arrs):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    ptrs = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return arrs, ptrs


def bonds(cores):
    return np.array([c.shape[0] for c in cores], dtype=np.int64)


def digit(x, L, j):
    return lib().or_digit(float(x), int(L), int(j))


def evaluate(parent, sdim, cores, vcoord, vdigit, Lc, xs, nthreads=None):
    xs = np.ascontiguousarray(np.atleast_2d(np.asarray(xs, dtype=np.float64)))
    if xs.shape[1] != len(Lc) and xs.shape[0] == len(Lc) and xs.shape[1] != 1:
        raise ValueError("xs must be [npts, ncoord]")
    npts = xs.shape[0]
    out = np.empty(npts)
    p, pp = _i(parent)
    s, sp = _i(sdim)
    b, bp = _l(bonds(cores))
    cs, cp = _dpp(cores)
    vc, vcp = _i(vcoord)
    vd, vdp = _i(vdigit)
    lc, lcp = _i(Lc)
    x, xp = _d(xs)
    o, op = _d(out)
    nt = nthreads or min(16, len(os.sched_getaffinity(0)))
    rc = lib().or_eval(len(p), pp, sp, bp, cp, vcp, vdp, len(lc), lcp, xp, npts, op, nt)
    return o, rc


def integrate(parent, sdim, cores, weight=0.5):
    p, pp = _i(parent)
    s, sp = _i(sdim)
    b, bp = _l(bonds(cores))
    cs, cp = _dpp(cores)
    out = C.c_double(0.0)
    rc = lib().or_integrate(len(p), pp, sp, bp, cp, float(weight), C.byref(out))
    if rc:
        raise RuntimeError(f"or_integrate failed: {rc}")
    return out.value


def inner(parent, sdim, coresA, coresB):
    p, pp = _i(parent)
    s, sp = _i(sdim)
    ba, bap = _l(bonds(coresA))
    bb, bbp = _l(bonds(coresB))
    ca, cap = _dpp(coresA)
    cb, cbp = _dpp(coresB)
    out = C.c_double(0.0)
    rc = lib().or_inner(len(p), pp, sp, bap, cap, bbp, cbp, C.byref(out))
    if rc:
        raise RuntimeError(f"or_inner failed: {rc}")
    return out.value


def _children(parent):
    ch = [[] for _ in parent]
    for v in range(1, len(parent)):
        ch[parent[v]].append(v)
    return ch


def contract(parent, sdim, coresW, coresA):
    """Exact operator application; returns the new cores."""
    ch = _children(parent)
    bw, ba = bonds(coresW), bonds(coresA)
    outs = []
    for v in range(len(parent)):
        shape = (bw[v] * ba[v], int(sdim[v])) + tuple(int(bw[c] * ba[c]) for c in ch[v])
        outs.append(np.zeros(shape))
    p, pp = _i(parent)
    s, sp = _i(sdim)
    b1, b1p = _l(bw)
    b2, b2p = _l(ba)
    cw, cwp = _dpp(coresW)
    ca, cap = _dpp(coresA)
    oo, op = _dpp(outs)
    rc = lib().or_contract(len(p), pp, sp, b1p, cwp, b2p, cap, op)
    if rc:
        raise RuntimeError(f"or_contract failed: {rc}")
    return oo


def contract_adjoint(parent, sdim, coresW, coresA, gbar, want_w=True):
    """Pullback of contract: (abar, wbar) for the cotangent gbar of its output."""
    bw, ba = bonds(coresW), bonds(coresA)
    abar = [np.zeros_like(np.asarray(c, dtype=np.float64)) for c in coresA]
    wbar = [np.zeros_like(np.asarray(c, dtype=np.float64)) for c in coresW] if want_w else None
    p, pp = _i(parent)
    s, sp = _i(sdim)
    b1, b1p = _l(bw)
    b2, b2p = _l(ba)
    cw, cwp = _dpp(coresW)
    ca, cap = _dpp(coresA)
    gg, gp = _dpp(gbar)
    aa, ap = _dpp(abar)
    if want_w:
        ww, wp = _dpp(wbar)
    rc = lib().or_contract_adjoint(len(p), pp, sp, b1p, cwp, b2p, cap, gp, ap, wp if want_w else None)
    if rc:
        raise RuntimeError(f"or_contract_adjoint failed: {rc}")
    return aa, (ww if want_w else None)


def truncate(parent, sdim, cores, cutoff=0.0, maxdim=1 << 30):
    """Returns (new cores, truncation error)."""
    ch = _children(parent)
    b = bonds(cores).copy()
    bufs = [np.ascontiguousarray(c, dtype=np.float64).ravel().copy() for c in cores]
    p, pp = _i(parent)
    s, sp = _i(sdim)
    bl, blp = _l(b)
    arrs, cp = _dpp(bufs)
    err = C.c_double(0.0)
    rc = lib().or_truncate(len(p), pp, sp, blp, cp, float(cutoff), int(maxdim), C.byref(err))
    if rc:
        raise RuntimeError(f"or_truncate failed: {rc}")
    out = []
    for v in range(len(p)):
        shape = (int(bl[v]), int(sdim[v])) + tuple(int(bl[c]) for c in ch[v])
        out.append(arrs[v][: int(np.prod(shape))].reshape(shape).copy())
    return out, err.value


def eval_grad(parent, sdim, cores, vcoord, vdigit, Lc, xs, ybar, nthreads=None):
    xs = np.ascontiguousarray(np.atleast_2d(np.asarray(xs, dtype=np.float64)))
    npts = xs.shape[0]
    grads = [np.zeros_like(np.asarray(c, dtype=np.float64)) for c in cores]
    p, pp = _i(parent)
    s, sp = _i(sdim)
    b, bp = _l(bonds(cores))
    cs, cp = _dpp(cores)
    vc, vcp = _i(vcoord)
    vd, vdp = _i(vdigit)
    lc, lcp = _i(Lc)
    x, xp = _d(xs)
    yb, ybp = _d(ybar)
    gg, gp = _dpp(grads)
    nt = nthreads or min(16, len(os.sched_getaffinity(0)))
    rc = lib().or_eval_grad(len(p), pp, sp, bp, cp, vcp, vdp, len(lc), lcp, xp, npts, ybp, gp, nt)
    if rc:
        raise RuntimeError(f"or_eval_grad failed: {rc}")
    return gg


def svd(M, cut=0.0):
    """Full SVD (cut = 0), or the factorization a truncation with relative
    cutoff cut starts from: rows below cut * ||M||_F^2 / k are not rotated
    against each other."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    r, c = M.shape
    k = min(r, c)
    U, S, Vt = np.zeros((r, k)), np.zeros(k), np.zeros((k, c))
    _, mp = _d(M)
    rc = lib().or_svd(r, c, mp, U.ctypes.data_as(C.POINTER(C.c_double)),
                      S.ctypes.data_as(C.POINTER(C.c_double)), Vt.ctypes.data_as(C.POINTER(C.c_double)),
                      float(cut))
    if rc:
        raise RuntimeError(f"or_svd failed: {rc}")
    return U, S, Vt


def qr(M):
    M = np.ascontiguousarray(M, dtype=np.float64)
    r, c = M.shape
    k = min(r, c)
    Q, R = np.zeros((r, k)), np.zeros((k, c))
    _, mp = _d(M)
    lib().or_qr(r, c, mp, Q.ctypes.data_as(C.POINTER(C.c_double)), R.ctypes.data_as(C.POINTER(C.c_double)))
    return Q, R


def trunc_rule(p, cutoff, maxdim):
    pa, pp = _d(p)
    err = C.c_double(0.0)
    m = lib().or_trunc_rule(pp, len(pa), float(cutoff), int(maxdim), C.byref(err))
    return int(m), err.value


# ---------------------------------------------------------------------------
# Function algebra on one layout, restated in numpy on the raw cores (the
# per-vertex Kronecker product of src/elementary_operators.jl:266-276 and the
# block direct sum of src/fixes.jl:55-95).
# ---------------------------------------------------------------------------
def hadamard(parent, sdim, coresG, coresF):
    """Cores of (g f)(x) = g(x) f(x): core v is the outer product of the two
    cores over every bond (g's index major), the site shared."""
    ch = _children(parent)
    out = []
    for v, (A, B) in enumerate(zip(coresG, coresF)):
        m = len(ch[v])
        la, lb = "abcdefgh"[:m], "ijklmnop"[:m]
        C = np.einsum(f"Ps{la},Qs{lb}->PQs" + "".join(x + y for x, y in zip(la, lb)), A, B)
        shape = (A.shape[0] * B.shape[0], A.shape[1]) + tuple(A.shape[2 + k] * B.shape[2 + k] for k in range(m))
        out.append(C.reshape(shape))
    return out


def direct_sum(parent, sdim, coresF, coresG):
    """Cores of (f + g)(x): every bond is the sum of the two, f's block first;
    the root's unit bond and the site are shared."""
    ch = _children(parent)
    out = []
    for v, (A, B) in enumerate(zip(coresF, coresG)):
        m = len(ch[v])
        pa, pb = A.shape[0], B.shape[0]
        p = 1 if v == 0 else pa + pb
        shape = (p, A.shape[1]) + tuple(A.shape[2 + k] + B.shape[2 + k] for k in range(m))
        C = np.zeros(shape)
        if v == 0 and m == 0:
            C[...] = A + B
        else:
            fa = (slice(0, 1) if v == 0 else slice(0, pa), slice(None)) + tuple(slice(0, A.shape[2 + k]) for k in range(m))
            fb = (slice(0, 1) if v == 0 else slice(pa, pa + pb), slice(None)) + tuple(
                slice(A.shape[2 + k], A.shape[2 + k] + B.shape[2 + k]) for k in range(m))
            C[fa] = A
            C[fb] = B
        out.append(C)
    return out


# ---------------------------------------------------------------------------
# Belief propagation on a network with loops, per point (the scalar of the
# projected network by ITensorNetworks' BP, as evaluate(f; alg="bp") uses it,
# src/itensornetworkfunction.jl:96-106), restated in numpy; and the exact
# scalar of a projected network by brute force, for small graphs.
# Cores: core v is [site][legs...], its legs being its edges in edge order.
# ---------------------------------------------------------------------------
def _project(nv, sdim, vcoord, vdigit, Lc, cores, x):
    T = []
    for v in range(nv):
        d = digit(x[vcoord[v]], int(Lc[vcoord[v]]), int(vdigit[v])) if vcoord[v] >= 0 else 0
        if d < 0:
            return None
        T.append(np.asarray(cores[v])[d])
    return T


def loopy_exact(nv, edges, sdim, vcoord, vdigit, Lc, cores, xs):
    """Exact scalar of every projected network (einsum over all edges)."""
    legs = [[e for e, (a, b) in enumerate(edges) if v in (a, b)] for v in range(nv)]
    out = []
    for x in np.atleast_2d(xs):
        T = _project(nv, sdim, vcoord, vdigit, Lc, cores, x)
        args = []
        for v in range(nv):
            args += [T[v], legs[v]]
        out.append(float(np.einsum(*args, [], optimize="greedy")))
    return np.array(out)


def bp_evaluate(nv, edges, sdim, vcoord, vdigit, Lc, cores, xs, maxiter=200, tol=1e-13):
    """BP estimate of every projected network: sweeps over the vertices in
    order, each vertex sending its messages over its legs in edge order,
    every message normalised to unit 2-norm and used at once; uniform start;
    stop when no entry of a sweep changes by more than tol or after maxiter
    sweeps; value prod_v Z_v / prod_e Z_e.  Returns (values, sweeps)."""
    legs = [[e for e, (a, b) in enumerate(edges) if v in (a, b)] for v in range(nv)]
    other = lambda e, v: edges[e][1] if edges[e][0] == v else edges[e][0]
    vals, sweeps = [], 0
    for x in np.atleast_2d(xs):
        T = _project(nv, sdim, vcoord, vdigit, Lc, cores, x)
        if T is None:
            vals.append(np.nan)
            continue
        # msg[(e, src)]: the message sent by src over edge e
        msg = {}
        for e, (a, b) in enumerate(edges):
            bd = T[a].shape[legs[a].index(e)]
            msg[(e, a)] = np.ones(bd)
            msg[(e, b)] = np.ones(bd)
        it = 0
        while it < maxiter:
            change = 0.0
            for v in range(nv):
                for k, e in enumerate(legs[v]):
                    args = [T[v], list(range(len(legs[v])))]
                    for kk, e2 in enumerate(legs[v]):
                        if kk != k:
                            args += [msg[(e2, other(e2, v))], [kk]]
                    m = np.einsum(*args, [k])
                    nrm = np.linalg.norm(m)
                    m = m / nrm if nrm > 0 else m * 0.0
                    change = max(change, float(np.max(np.abs(m - msg[(e, v)]))))
                    msg[(e, v)] = m
            it += 1
            if change <= tol:
                break
        sweeps = max(sweeps, it)
        val = 1.0
        for v in range(nv):
            args = [T[v], list(range(len(legs[v])))]
            for kk, e2 in enumerate(legs[v]):
                args += [msg[(e2, other(e2, v))], [kk]]
            val *= float(np.einsum(*args, []))
        for e, (a, b) in enumerate(edges):
            val /= float(msg[(e, a)] @ msg[(e, b)])
        vals.append(val)
    return np.array(vals), sweeps

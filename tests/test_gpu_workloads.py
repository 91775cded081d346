"""Engine against the oracle at the benchmark workloads' own sizes.

- Fredholm batch: the costliest members of the 512-chain batch (bond 285
  after the operator, cores far beyond LDS, so the batch launch takes the
  global-workspace moves without helper workgroups) together with members
  just above and below the LDS limit, through one operate_batch call.
- 2D fit: the gradient of the MSE over 2^20 points at bond 32, against the
  oracle's adjoint on a subset, against central differences of the loss on a
  few entries, and split-sum linearity over the full point set.
- 2D Laplacian: the truncated result evaluated at all 1e6 points.
Bars as in test_gpu_parity: 1e-12 relative for float64 values of O(1)
results, 1e-11 norm-relative for truncated operator results.
"""
import ctypes
import math

import numpy as np
import pytest

from helpers import oracle_eval, rel_l2

pytestmark = pytest.mark.gpu

LDS_DOUBLES = 20224  # QTT_LDS_DOUBLES: cores above this take the global workspace


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _max_core(w):
    b = w.f.bonds() * w.op.bonds()
    nv = w.f.topo.nv
    return max(int(b[v]) * 2 * (int(b[v + 1]) if v + 1 < nv else 1) for v in range(nv))


def _cost(w):
    b = w.f.bonds() * w.op.bonds()
    nv = w.f.topo.nv
    return sum(int(b[v]) * 2 * (int(b[v + 1]) if v + 1 < nv else 1) * int(b[v]) for v in range(nv))


def test_fredholm_batch_real_bond_range(q, po, eng, rng):
    import torch
    from itensornumericalanalysis_jl_amd import workloads
    ws = workloads.fredholm_batch(npts=4096)
    order = sorted(range(len(ws)), key=lambda i: -_cost(ws[i]))
    above = [i for i in order if _max_core(ws[i]) > LDS_DOUBLES]
    below = [i for i in order if _max_core(ws[i]) <= LDS_DOUBLES]
    pick = order[:8] + above[-4:] + below[:2] + below[-2:]
    assert len(set(pick)) == len(pick) == 16
    assert sum(_max_core(ws[i]) > LDS_DOUBLES for i in pick) >= 12
    sel = [ws[i] for i in pick]
    print("members", pick, "largest core", max(_max_core(w) for w in sel))
    gs, errs = q.operate_batch([w.op for w in sel], [w.f for w in sel], cutoff=1e-10)
    ints = q.integrate_batch(gs)
    ys = q.evaluate_batch(gs, [torch.tensor(w.xs, device="cuda") for w in sel])
    eng.sync()
    for w, g, e, integ, y in zip(sel, gs, errs, ints, ys):
        t = w.f.topo
        C = po.contract(t.parent, t.sdim, w.op.cores, w.f.cores)
        gref, rerr = po.truncate(t.parent, t.sdim, C, 1e-10)
        assert list(g.bonds()) == [c.shape[0] for c in gref], w.name
        assert e == pytest.approx(rerr, rel=1e-6, abs=1e-18), w.name
        iref = po.integrate(t.parent, t.sdim, gref, 0.5)
        assert integ == pytest.approx(iref, rel=1e-11, abs=1e-11 * np.abs(iref)), w.name
        yr, rc = po.evaluate(t.parent, t.sdim, gref, t.vcoord, t.vdigit, t.Lc, w.xs)
        assert rc == 0
        yy = y.cpu().numpy()
        assert np.max(np.abs(yy - yr)) <= 1e-11 * np.max(np.abs(yr)), w.name


def test_fit_2d_gradient_full_size(q, po, eng):
    import ctypes
    import torch
    from itensornumericalanalysis_jl_amd import _capi, workloads
    w = workloads.fit_2d()
    f = w.f
    n = len(w.xs)
    assert n == 1 << 20 and f.maxlinkdim() == 32
    xs = torch.tensor(w.xs, device="cuda")
    target = torch.tensor(w.extra["target"], device="cuda")
    y = q.evaluate(f, xs)
    ybar = (2.0 / n) * (y - target)  # d MSE / d f(x_m)
    dn = q.api.device_network(f)
    ctx = eng

    def grad_of(xd, yb):
        out = torch.empty(dn.size(), dtype=torch.float64, device="cuda")
        ctx.check(_capi.lib().qtt_eval_grad(ctx.handle, dn.handle, ctypes.c_void_p(xd.data_ptr()),
                                            ctypes.c_void_p(yb.data_ptr()), xd.shape[0],
                                            ctypes.c_void_p(out.data_ptr()), 0))
        ctx.sync()
        return out

    G = grad_of(xs, ybar)
    # linearity over the points: the two halves' gradients add up to the whole
    h = n // 2
    G2 = grad_of(xs[:h].contiguous(), ybar[:h].contiguous()) + grad_of(xs[h:].contiguous(), ybar[h:].contiguous())
    assert (torch.linalg.norm(G - G2) / torch.linalg.norm(G)).item() < 1e-13
    # the oracle's adjoint on a subset of 2^14 points
    ns = 1 << 14
    sub = q.eval_grad(f, w.xs[:ns], ybar[:ns].cpu().numpy())
    t = f.topo
    ref = po.eval_grad(t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc, w.xs[:ns], ybar[:ns].cpu().numpy())
    num = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(sub, ref)))
    den = math.sqrt(sum(float(np.sum(b ** 2)) for b in ref))
    assert num / den < 1e-12
    # central differences (h = 1e-6) over all 2^20 points of S(A) = sum_m
    # ybar_m f_A(x_m), with ybar the MSE cotangent above: its gradient is the
    # gradient of the loss (chain rule), and S is linear in any single core
    # entry, so the difference quotient is exact up to the rounding of S, taken
    # here as 1e-12 relative per value: 1e-12 * sum |ybar_m f(x_m)| / h.
    # (Differencing the loss itself would drown these entries, about 2e-11 at
    # this initialisation, in the rounding of a loss of 0.05.)
    Gh = G.cpu().numpy()
    offs = np.cumsum([0] + [c.size for c in f.cores])
    cores = [c.copy() for c in f.cores]
    scale = float(torch.sum(torch.abs(ybar * y)).item())
    rounding = 1e-12 * scale / 1e-6
    # the two largest entries and two random ones
    picks = list(np.argsort(-np.abs(Gh))[:2]) + list(np.random.default_rng(7).integers(0, Gh.size, 2))
    for flat in picks:
        v = int(np.searchsorted(offs, flat, side="right") - 1)
        e = int(flat - offs[v])
        vals = []
        for sgn in (1.0, -1.0):
            pert = [c.copy() for c in cores]
            pert[v].flat[e] += sgn * 1e-6
            fp = q.ITensorNetworkFunction(f.s, pert, f.topo)
            vals.append(float(torch.dot(ybar, q.evaluate(fp, xs)).item()))
        fd = (vals[0] - vals[1]) / 2e-6
        print(f"core {v} entry {e}: adjoint {Gh[flat]:.12e} central difference {fd:.12e}")
        assert abs(fd - Gh[flat]) <= 1e-6 * abs(Gh[flat]) + rounding, (v, e)
    assert abs(Gh[picks[0]]) > 1e3 * rounding  # the check has teeth


def test_laplacian_2d_result_at_1e6_points(q, po, eng):
    import torch
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.laplacian_2d()
    fc = q.truncate(w.f, cutoff=1e-16, maxdim=w.extra["compress_maxdim"])
    g = q.operate(w.op, fc, cutoff=w.cutoff, maxdim=w.maxdim)
    xd = torch.tensor(w.xs, device="cuda")
    y = q.evaluate(g, xd).cpu().numpy()
    eng.sync()
    t = g.topo
    yr, rc = po.evaluate(t.parent, t.sdim, g.cores, t.vcoord, t.vdigit, t.Lc, w.xs, nthreads=8)
    assert rc == 0
    # the result is dominated by boundary layers: compare relative to its scale
    assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))
    assert rel_l2(y, yr) < 1e-12


# ---------------------------------------------------------------- boundary behaviour
def test_truncate_noise_block_finished(q, po, eng):
    # rows below the Jacobi noise level whose joint singular value outranks a
    # kept row: the engine must keep what a full SVD keeps (LAPACK)
    from test_oracle import noise_block_network
    from itensornumericalanalysis_jl_amd import _capi
    topo, cores, dense_ref = noise_block_network()
    dn = _capi.DeviceNetwork.upload(eng, topo, cores)
    err = ctypes.c_double(0.0)
    eng.check(_capi.lib().qtt_truncate(eng.handle, dn.handle, 0.1, 0, ctypes.byref(err)))
    assert list(dn.bonds()) == [1, 2]
    assert err.value == pytest.approx(0.045, rel=1e-12)
    out = dn.download()
    assert np.allclose(out[0].reshape(8, 2) @ out[1], dense_ref, atol=1e-15)


def _normrel(q, po, g, gref, f, rng, npts=4000):
    """max |g - gref| / max |gref| over random points and the grid's edge lines
    (the boundary layers of operator results live there)."""
    nd = f.dimension()
    xs = [rng.uniform(0, 1, (npts, nd))]
    for d in range(nd):
        h = 2.0 ** -int(f.topo.Lc[d])
        for val in (0.0, h, 1 - h):
            e = rng.uniform(0, 1, (64, nd))
            e[:, d] = val
            xs.append(e)
    xs = np.concatenate(xs)
    y = q.evaluate(g, xs)
    t = f.topo
    yr, _ = po.evaluate(t.parent, t.sdim, gref, t.vcoord, t.vdigit, t.Lc, xs)
    return float(np.max(np.abs(y - yr)) / np.max(np.abs(yr)))


@pytest.mark.parametrize("case", ["laplacian_2d", "comb_global", "fredholm_members"])
def test_operate_mixed_precision(q, po, eng, rng, case):
    # float-float (F32_SPLIT) bulk products meet 1e-10 against the float64
    # oracle; plain float32 (F32) is about 1e-7.  QR, SVD and the cutoff run in
    # float64 either way.
    from itensornumericalanalysis_jl_amd import workloads
    P = q.Precision
    if case == "laplacian_2d":
        w = workloads.laplacian_2d(npts=10)
        t = w.f.topo
        fref, _ = po.truncate(t.parent, t.sdim, w.f.cores, 1e-16, 64)
        f = q.ITensorNetworkFunction(w.s, fref, t)
        items = [(w.op, f, w.cutoff, w.maxdim)]
    elif case == "comb_global":  # moves beyond LDS: global workspace, helper workgroups
        w = workloads.comb_3d(bond=40)
        items = [(w.op, w.f, w.cutoff, w.maxdim)]
    else:
        ws = workloads.fredholm_batch(npts=16)
        big = sorted(range(len(ws)), key=lambda i: -ws[i].f.maxlinkdim())[:2]
        items = [(ws[i].op, ws[i].f, 1e-10, 0) for i in big] + [(ws[3].op, ws[3].f, 1e-10, 0)]
    for op, f, cut, md in items:
        t = f.topo
        C = po.contract(t.parent, t.sdim, op.cores, f.cores)
        gref, _ = po.truncate(t.parent, t.sdim, C, cut, md or (1 << 30))
        errs = {}
        for prec in (P.F32_SPLIT, P.F32):
            g = q.operate(op, f, cutoff=cut, maxdim=md or None, precision=prec)
            errs[prec] = _normrel(q, po, g, gref, f, rng)
        if case == "fredholm_members":  # the batch launch too
            gb, _ = q.operate_batch([op], [f], cutoff=cut, precision=P.F32_SPLIT)
            errs["batch_split"] = _normrel(q, po, gb[0], gref, f, rng)
        print(f"{case}: max-norm error vs float64 oracle: split {errs[P.F32_SPLIT]:.2e}, "
              f"float32 {errs[P.F32]:.2e}")
        assert errs[P.F32_SPLIT] < 1e-10
        assert errs.get("batch_split", 0.0) < 1e-10
        assert 1e-12 < errs[P.F32] < 1e-4


def test_async_status_not_reported_by_sync_calls(q, eng):
    # a NaN point through the asynchronous (device tensor) path leaves its
    # status for qtt_sync; synchronous calls in between must not report it
    import torch
    s = q.continuous_siteinds(q.named_grid((8, 1)))
    f = q.exp_itn(s)
    bad = torch.tensor([[0.25], [float("nan")]], dtype=torch.float64, device="cuda")
    y = q.evaluate(f, bad)
    assert q.evaluate(f, 0.5) == pytest.approx(math.exp(0.5), rel=1e-13)
    g = q.eval_grad(f, np.array([[0.5]]), np.array([1.0]))
    assert len(g) == f.topo.nv
    assert math.isnan(y[1].item())
    from itensornumericalanalysis_jl_amd._capi import EngineError
    with pytest.raises(EngineError) as ei:
        eng.sync()
    assert ei.value.code == 2  # QTT_ERANGE
    eng.sync()  # cleared


def test_evaluate_on_torch_stream(q, po, eng, rng):
    # under `with torch.cuda.stream(s)` the engine's reads and writes are
    # ordered with torch's work on s
    import torch
    s = q.continuous_siteinds(q.named_grid((16, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=8, rng=rng, std=0.4)
    xs = rng.uniform(0, 1, (300000, 2))
    ref = oracle_eval(po, f, xs)
    prev = eng.stream
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        xd = torch.tensor(xs, device="cuda") * 1.0  # produced on st
        y = q.evaluate(f, xd)
        y2 = y * 2.0  # consumed on st
    st.synchronize()
    assert rel_l2(y2.cpu().numpy() / 2.0, ref) < 1e-12
    assert eng.stream == prev  # the context's stream is restored


# ---------------------------------------------------------------- status paths
def _raw_topo(parent, sdim, vcoord, vdigit, ndigits):
    from itensornumericalanalysis_jl_amd import _capi
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in (parent, sdim, vcoord, vdigit, ndigits)]
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    st = _capi.QttTopology(len(keep[0]), ip(keep[0]), ip(keep[1]), ip(keep[2]), ip(keep[3]), len(keep[4]),
                           ip(keep[4]))
    return st, keep


def _create(eng, parent, sdim, vcoord, vdigit, ndigits, bond):
    """qtt_net_create with zero cores of the implied shapes; returns (rc, handle)."""
    from itensornumericalanalysis_jl_amd import _capi
    st, keep = _raw_topo(parent, sdim, vcoord, vdigit, ndigits)
    nv = len(parent)
    ch = [[c for c in range(nv) if parent[c] == v] for v in range(nv)]
    cores = [np.zeros(max(1, int(np.prod([bond[v], sdim[v]] + [bond[c] for c in ch[v]])))) for v in range(nv)]
    b = np.ascontiguousarray(bond, dtype=np.int64)
    h = ctypes.c_void_p()
    rc = _capi.lib().qtt_net_create(eng.handle, ctypes.byref(st), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    _capi._dpp(cores), ctypes.byref(h))
    return rc, h


def test_status_einval_paths(q, eng):
    from itensornumericalanalysis_jl_amd import _capi
    L = _capi.lib()
    ok = ([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3], [1, 2, 2])
    rc, h = _create(eng, *ok)
    assert rc == 0
    L.qtt_net_destroy(h)
    bad = {
        "root parent": ([0, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3], [1, 2, 2]),
        "parent after child": ([-1, 2, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3], [1, 2, 2]),
        "root bond": ([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3], [2, 2, 2]),
        "duplicate digit": ([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 2], [3], [1, 2, 2]),
        "digit out of range": ([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 4], [3], [1, 2, 2]),
        "non-binary digit site": ([-1, 0, 1], [2, 3, 2], [0, 0, 0], [1, 2, 3], [3], [1, 2, 2]),
        "zero bond": ([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3], [1, 0, 2]),
    }
    for name, args in bad.items():
        rc, h = _create(eng, *args)
        assert rc == _capi.QTT_EINVAL, name
        assert L.qtt_last_error(eng.handle), name
    # calls on mismatched layouts, bad sizes and unknown precisions
    s3 = q.continuous_siteinds(q.named_grid((3, 1)))
    s4 = q.continuous_siteinds(q.named_grid((4, 1)))
    f3, f4 = q.exp_itn(s3), q.exp_itn(s4)
    with pytest.raises(_capi.EngineError) as ei:
        q.inner(f3, f4)
    assert ei.value.code == _capi.QTT_EINVAL
    with pytest.raises(_capi.EngineError) as ei:
        q.operate(q.second_derivative_operator(s4), f3, cutoff=1e-12)
    assert ei.value.code == _capi.QTT_EINVAL
    d3 = q.api.device_network(f3)
    err = ctypes.c_double()
    assert L.qtt_truncate(eng.handle, d3.handle, 0.0, -1, ctypes.byref(err)) == _capi.QTT_EINVAL
    out = np.zeros(4)
    xs = np.zeros((4, 1))
    assert L.qtt_eval_host(eng.handle, d3.handle, _capi._dptr(xs), -1, _capi._dptr(out), 0) == _capi.QTT_EINVAL
    assert L.qtt_eval_host(eng.handle, d3.handle, _capi._dptr(xs), 4, _capi._dptr(out), 7) == _capi.QTT_EINVAL


def test_status_cutoff_and_maxdim_semantics(q, eng, rng):
    s = q.continuous_siteinds(q.named_grid((10, 1)))
    f = q.random_itn(s, link_space=6, rng=rng)
    a, ea = q.truncate(f, cutoff=-1.0, return_error=True)  # a negative cutoff truncates nothing but round-off
    b, eb = q.truncate(f, cutoff=0.0, return_error=True)
    assert list(a.bonds()) == list(b.bonds()) and ea == eb
    c = q.truncate(f, cutoff=0.0, maxdim=0)  # maxdim 0: no limit
    d = q.truncate(f, cutoff=0.0, maxdim=1 << 30)
    assert list(c.bonds()) == list(d.bonds())
    assert max(q.truncate(f, cutoff=0.0, maxdim=3).bonds()) == 3
    op = q.second_derivative_operator(s)
    g = q.operate(op, f, truncate=False)  # cutoff < 0 at the C-ABI: the exact product
    assert list(g.bonds()) == list(f.bonds() * op.bonds())


def test_status_enomem(q, eng):
    # bond 2^20 on both edges of a three-site chain: the middle core alone is
    # 2^41 doubles (16 TiB), so the allocation fails before any core is read
    from itensornumericalanalysis_jl_amd import _capi
    st, keep = _raw_topo([-1, 0, 1], [2, 2, 2], [0, 0, 0], [1, 2, 3], [3])
    b = np.array([1, 1 << 20, 1 << 20], dtype=np.int64)
    dummy = [np.zeros(1) for _ in range(3)]
    h = ctypes.c_void_p()
    rc = _capi.lib().qtt_net_create(eng.handle, ctypes.byref(st), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    _capi._dpp(dummy), ctypes.byref(h))
    assert rc == _capi.QTT_ENOMEM
    assert not h.value
    # the context still works
    s = q.continuous_siteinds(q.named_grid((4, 1)))
    assert q.evaluate(q.exp_itn(s), 0.5) == pytest.approx(math.exp(0.5), rel=1e-13)


def test_status_enoconv(q, eng, rng, monkeypatch):
    # a context whose Jacobi SVD may take one sweep only: a random bond-12
    # chain needs more, so the truncation reports QTT_ENOCONV
    from itensornumericalanalysis_jl_amd import _capi
    monkeypatch.setenv("QTT_JACOBI_MAX_SWEEPS", "1")
    ctx = _capi.Context(0)
    try:
        s = q.continuous_siteinds(q.named_grid((8, 1)))
        f = q.random_itn(s, link_space=12, rng=rng)
        dn = _capi.DeviceNetwork.upload(ctx, f.topo, f.cores)
        err = ctypes.c_double()
        rc = _capi.lib().qtt_truncate(ctx.handle, dn.handle, 1e-4, 0, ctypes.byref(err))
        assert rc == _capi.QTT_ENOCONV
        dn._release()
    finally:
        ctx.close()
    # the process-wide context is unaffected
    g = q.truncate(f, cutoff=1e-4)
    assert max(g.bonds()) <= 12

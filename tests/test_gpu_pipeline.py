"""The batch pipeline (qtt_batch_run): contract -> truncate -> integrate ->
evaluate of many functions as one HIP graph, against the same stages run as
separate batch calls (same kernels, so bonds, errors, integrals and values
agree to the bit) and against the oracle.

- Fredholm batch members with cores beyond LDS and small ones, 4096 points each;
- a batch mixing chains whose final bonds exceed the register kernels' 32 and
  a comb tree (both evaluated by the single-network path after the graph);
- no truncation (cutoff < 0), no points, an invalid point (ERANGE, outputs
  destroyed) and a context that is usable afterwards.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LDS_DOUBLES = 20224


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


def _separate(q, eng, ops, fs, xs, cutoff, maxdim=None):
    gs, errs = q.operate_batch(ops, fs, cutoff=cutoff, maxdim=maxdim)
    ints = q.integrate_batch(gs)
    ys = q.evaluate_batch(gs, xs)
    eng.sync()
    return gs, errs, ints, ys


def test_pipeline_fredholm_members(q, po, eng):
    import torch
    from itensornumericalanalysis_jl_amd import workloads
    ws = workloads.fredholm_batch(npts=4096)
    above = [i for i in range(len(ws)) if _max_core(ws[i]) > LDS_DOUBLES]
    below = [i for i in range(len(ws)) if _max_core(ws[i]) <= LDS_DOUBLES]
    pick = above[:6] + below[:10]
    sel = [ws[i] for i in pick]
    ops, fs = [w.op for w in sel], [w.f for w in sel]
    xs = [torch.tensor(w.xs, device="cuda") for w in sel]
    print("pipeline members", pick, flush=True)
    gs, errs, ints, ys = q.operate_batch_pipeline(ops, fs, xs, cutoff=1e-10)
    gs2, errs2, ints2, ys2 = _separate(q, eng, ops, fs, xs, 1e-10)
    for k, w in enumerate(sel):
        assert list(gs[k].bonds()) == list(gs2[k].bonds()), w.name
        assert errs[k] == errs2[k], w.name
        assert ints[k] == ints2[k], w.name
        assert torch.equal(ys[k], ys2[k]), w.name
    # the oracle on two members above and two below the LDS limit
    for k in (0, 1, 6, 7):
        w = sel[k]
        t = w.f.topo
        C = po.contract(t.parent, t.sdim, w.op.cores, w.f.cores)
        gref, rerr = po.truncate(t.parent, t.sdim, C, 1e-10)
        assert list(gs[k].bonds()) == [c.shape[0] for c in gref], w.name
        assert errs[k] == pytest.approx(rerr, rel=1e-6, abs=1e-18), w.name
        iref = po.integrate(t.parent, t.sdim, gref, 0.5)
        assert ints[k] == pytest.approx(iref, rel=1e-11, abs=1e-11 * abs(iref)), w.name
        yr, rc = po.evaluate(t.parent, t.sdim, gref, t.vcoord, t.vdigit, t.Lc, w.xs)
        assert rc == 0
        assert np.max(np.abs(ys[k].cpu().numpy() - yr)) <= 1e-11 * np.max(np.abs(yr)), w.name


def test_pipeline_wide_bonds_and_tree(q, po, eng, rng):
    """Final bonds above 32 and a comb: evaluated after the graph, one launch each."""
    import torch
    from itensornumericalanalysis_jl_amd import workloads
    fs, ops, xs = [], [], []
    for L, chi in ((12, 48), (10, 6), (14, 40)):
        s = q.continuous_siteinds(q.named_grid((L, 1)))
        fs.append(q.random_itn(s, link_space=chi, rng=rng))
        # a shift keeps a random function's ranks; the kernel operator smooths them
        ops.append(q.forward_shift_op(s) if chi > 32 else q.fredholm_operator(s))
        xs.append(rng.uniform(0.0, 1.0, (777, 1)))
    w = workloads.comb_3d(bond=8, digits=6)
    fs.append(w.f)
    ops.append(w.op)
    xs.append(rng.uniform(0.0, 1.0, (555, w.f.dimension())))
    xt = [torch.tensor(x, device="cuda") for x in xs]
    gs, errs, ints, ys = q.operate_batch_pipeline(ops, fs, xt, cutoff=1e-12)
    assert max(gs[0].bonds()) > 32 and max(gs[2].bonds()) > 32  # the single-network path ran
    gs2, errs2, ints2, ys2 = _separate(q, eng, ops, fs, xt, 1e-12)
    for k in range(len(fs)):
        assert list(gs[k].bonds()) == list(gs2[k].bonds())
        assert errs[k] == errs2[k]
        assert ints[k] == ints2[k]
        assert torch.equal(ys[k], ys2[k])
        t = fs[k].topo
        C = po.contract(t.parent, t.sdim, ops[k].cores, fs[k].cores)
        gref, _ = po.truncate(t.parent, t.sdim, C, 1e-12)
        yr, rc = po.evaluate(t.parent, t.sdim, gref, t.vcoord, t.vdigit, t.Lc, xs[k])
        assert rc == 0
        assert np.max(np.abs(ys[k].cpu().numpy() - yr)) <= 1e-10 * np.max(np.abs(yr))


def test_pipeline_exact_and_without_points(q, po, eng, rng):
    import torch
    fs, ops = [], []
    for L, chi in ((9, 5), (11, 12), (8, 3)):
        s = q.continuous_siteinds(q.named_grid((L, 1)))
        fs.append(q.random_itn(s, link_space=chi, rng=rng))
        ops.append(q.fredholm_operator(s))
    # no truncation: bonds multiply, errors are zero
    xs = [torch.tensor(rng.uniform(0.0, 1.0, (300, 1)), device="cuda") for _ in fs]
    gs, errs, ints, ys = q.operate_batch_pipeline(ops, fs, xs, truncate=False)
    for k, f in enumerate(fs):
        assert list(gs[k].bonds()) == list(f.bonds() * ops[k].bonds())
        assert errs[k] == 0.0
        t = f.topo
        C = po.contract(t.parent, t.sdim, ops[k].cores, f.cores)
        assert ints[k] == pytest.approx(po.integrate(t.parent, t.sdim, C, 0.5), rel=1e-12)
        yr, rc = po.evaluate(t.parent, t.sdim, C, t.vcoord, t.vdigit, t.Lc, xs[k].cpu().numpy())
        assert rc == 0
        assert np.max(np.abs(ys[k].cpu().numpy() - yr)) <= 1e-12 * np.max(np.abs(yr))
    # without points: operate and integrate only
    gs, errs, ints, ys = q.operate_batch_pipeline(ops, fs, None, cutoff=1e-10, maxdim=4)
    assert ys is None
    for k, f in enumerate(fs):
        assert max(gs[k].bonds()) <= 4
        g1 = q.operate(ops[k], f, cutoff=1e-10, maxdim=4)
        assert list(g1.bonds()) == list(gs[k].bonds())
        assert ints[k] == pytest.approx(q.integrate(g1), rel=1e-13, abs=1e-300)


def test_pipeline_invalid_point(q, eng, rng):
    import torch
    s = q.continuous_siteinds(q.named_grid((10, 1)))
    f = q.random_itn(s, link_space=4, rng=rng)
    op = q.fredholm_operator(s)
    x = rng.uniform(0.0, 1.0, (100, 1))
    x[37, 0] = float("nan")
    with pytest.raises(q.DigitMapError):
        q.operate_batch_pipeline([op], [f], [torch.tensor(x, device="cuda")], cutoff=1e-10)
    # the context stays usable and the status was consumed
    x[37, 0] = 0.5
    gs, _, _, ys = q.operate_batch_pipeline([op], [f], [torch.tensor(x, device="cuda")], cutoff=1e-10)
    g1 = q.operate(op, f, cutoff=1e-10)
    yr = q.evaluate(g1, x)
    assert np.max(np.abs(ys[0].cpu().numpy() - yr)) <= 1e-12 * np.max(np.abs(yr))

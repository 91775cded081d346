"""Engine (HIP, through the C-ABI) against the CPU oracle on identical inputs.

Bars: float64 results within 1e-12 relative of the oracle (norm-wise over the
points, and per value where the values are O(1)); integer digit work exact
(checked through values at grid points and the invalid-coordinate status).
"""
import math

import numpy as np
import pytest

from helpers import grid_of, oracle_eval, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _oracle_apply_truncate(po, f, op, cutoff, maxdim):
    t = f.topo
    C = po.contract(t.parent, t.sdim, op.cores, f.cores)
    if cutoff is None or cutoff < 0:
        return C, 0.0
    return po.truncate(t.parent, t.sdim, C, cutoff, maxdim or (1 << 30))


# ------------------------------------------------------------------ evaluate
def test_eval_sine_1d(q, po, eng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.sine_1d(npts=20000)
    y = q.evaluate(w.f, w.xs)
    ref = oracle_eval(po, w.f, w.xs)
    assert rel_l2(y, ref) < TOL
    assert np.max(np.abs(y - ref)) < TOL * np.max(np.abs(ref))
    # and the closed form on the grid
    xg = grid_of(w.xs[:, 0], 12)
    assert np.allclose(y, workloads.sine_1d_function(xg), rtol=1e-12, atol=1e-12)


def test_eval_laplacian_2d(q, po, eng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.laplacian_2d(npts=20000)
    y = q.evaluate(w.f, w.xs)
    ref = oracle_eval(po, w.f, w.xs)
    assert rel_l2(y, ref) < TOL
    xg = grid_of(w.xs, 20)
    closed = workloads.laplacian_2d_function(xg, w.extra["a"], w.extra["w"], w.extra["b"])
    assert rel_l2(y, closed) < 1e-12


@pytest.mark.parametrize("bond", [1, 3, 8, 17, 32, 64])
def test_eval_random_chain_bonds(q, po, eng, rng, bond):
    s = q.continuous_siteinds(q.named_grid((18, 1)), map_dimension=3)
    f = q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond))
    xs = rng.uniform(0, 1, (3001, 3))
    assert rel_l2(q.evaluate(f, xs), oracle_eval(po, f, xs)) < TOL


@pytest.mark.parametrize("bond", [5, 16])
def test_eval_many_points(q, po, eng, rng, bond):
    # >= 2^19 points with bonds <= 16 take the two-points-per-lane kernel; a ragged count
    s = q.continuous_siteinds(q.named_grid((16, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond))
    xs = rng.uniform(0, 1, ((1 << 19) + 123, 2))
    y = q.evaluate(f, xs)
    ref = oracle_eval(po, f, xs)
    assert rel_l2(y, ref) < TOL
    # a small call takes the plain chain (no environment tables): same values up to rounding
    assert np.allclose(y[:4096], q.evaluate(f, xs[:4096]), rtol=1e-13, atol=1e-13 * np.abs(y).max())


def _eval_with_tables(q, f, xs, tbits, monkeypatch):
    """f at xs through a context whose chain evaluate takes environment tables
    of tbits digit sites (clamped to half of the chain's digits); returns the
    values and the status of the following qtt_sync."""
    import ctypes
    import torch
    from itensornumericalanalysis_jl_amd import _capi
    monkeypatch.setenv("QTT_EVAL_TABLE_BITS", str(tbits))
    ctx = _capi.Context(0)
    monkeypatch.delenv("QTT_EVAL_TABLE_BITS")
    try:
        dn = _capi.DeviceNetwork.upload(ctx, f.topo, f.cores)
        xd = torch.tensor(xs, device="cuda")
        out = torch.empty(len(xs), dtype=torch.float64, device="cuda")
        ctx.check(_capi.lib().qtt_eval(ctx.handle, dn.handle, ctypes.c_void_p(xd.data_ptr()), len(xs),
                                       ctypes.c_void_p(out.data_ptr()), 0))
        rc = _capi.lib().qtt_sync(ctx.handle)
        y = out.cpu().numpy()
        dn._release()
        return y, rc
    finally:
        ctx.close()


@pytest.mark.parametrize("tbits", [1, 3, 8, 64])
@pytest.mark.parametrize("case", ["2d", "3d_odd", "no_digit"])
def test_eval_environment_tables(q, po, eng, rng, monkeypatch, tbits, case):
    # f = L[prefix digits] (middle sites) R[suffix digits]: tables of 1, 3 and 8
    # digit sites, and as many as allowed (half of the digits each: no middle
    # site on the even chain, one on the odd one); vertices without a digit
    # inside the tables and in the middle
    if case == "2d":
        s = q.continuous_siteinds(q.named_grid((20, 1)), map_dimension=2)
        bond = 16
    elif case == "3d_odd":
        s = q.continuous_siteinds(q.named_grid((21, 1)), map_dimension=3)
        bond = 7
    else:
        g = q.named_grid((15, 1))
        s = q.continuous_siteinds(g, [[(j, 1) for j in range(2, 8)], [(j, 1) for j in range(9, 15)]])
        bond = 5
    f = q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond))
    nd = f.dimension()
    xs = rng.uniform(0, 1, (40000, nd))
    xs[0] = 0.0
    xs[1] = np.nextafter(1.0, 0.0)
    xs[2] = 1.0
    y, rc = _eval_with_tables(q, f, xs, tbits, monkeypatch)
    assert rc == 0
    ref = oracle_eval(po, f, xs)
    assert rel_l2(y, ref) < TOL
    assert np.max(np.abs(y - ref)) < TOL * np.max(np.abs(ref))


def test_eval_tables_kept_until_cores_change(q, po, eng, rng, monkeypatch):
    """The tables built by one evaluate serve the next ones (other points, the
    same table size), are rebuilt for another size, and are dropped when a
    truncation changes the cores in place."""
    import ctypes
    import torch
    from itensornumericalanalysis_jl_amd import _capi
    s = q.continuous_siteinds(q.named_grid((20, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=12, rng=rng, std=1.0 / math.sqrt(12))
    monkeypatch.setenv("QTT_EVAL_TABLE_BITS", "6")
    ctx = _capi.Context(0)
    monkeypatch.delenv("QTT_EVAL_TABLE_BITS")
    try:
        dn = _capi.DeviceNetwork.upload(ctx, f.topo, f.cores)

        def ev(xs):
            xd = torch.tensor(xs, device="cuda")
            out = torch.empty(len(xs), dtype=torch.float64, device="cuda")
            ctx.check(_capi.lib().qtt_eval(ctx.handle, dn.handle, ctypes.c_void_p(xd.data_ptr()), len(xs),
                                           ctypes.c_void_p(out.data_ptr()), 0))
            ctx.check(_capi.lib().qtt_sync(ctx.handle))
            return out.cpu().numpy()

        for _ in range(3):  # first call builds the tables, the others reuse them
            xs = rng.uniform(0, 1, (5000, 2))
            assert rel_l2(ev(xs), oracle_eval(po, f, xs)) < TOL
        # truncate in place: the cores (and bonds) change under the same handle
        err = ctypes.c_double(0.0)
        ctx.check(_capi.lib().qtt_truncate(ctx.handle, dn.handle, 0.0, 5, ctypes.byref(err)))
        assert err.value > 0.0
        cores = dn.download()
        t = f.topo
        xs = rng.uniform(0, 1, (5000, 2))
        ref, rc = po.evaluate(t.parent, t.sdim, cores, t.vcoord, t.vdigit, t.Lc, xs)
        assert rc == 0
        assert rel_l2(ev(xs), ref) < TOL
        dn._release()
    finally:
        ctx.close()


def test_eval_environment_tables_invalid_point(q, eng, rng, monkeypatch):
    s = q.continuous_siteinds(q.named_grid((16, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=6, rng=rng, std=0.4)
    xs = rng.uniform(0, 1, (50000, 2))
    xs[123, 1] = float("nan")
    xs[456, 0] = -0.5
    y, rc = _eval_with_tables(q, f, xs, 6, monkeypatch)
    assert rc == 2  # QTT_ERANGE at the next sync
    assert np.isnan(y[123]) and np.isnan(y[456])
    assert np.isfinite(np.delete(y, [123, 456])).all()


def test_eval_tree(q, po, eng, rng):
    s = q.continuous_siteinds(q.named_comb_tree((3, 5)), map_dimension=2)
    f = q.random_itn(s, link_space=4, rng=rng, std=0.5)
    xs = rng.uniform(0, 1, (2000, 2))
    assert rel_l2(q.evaluate(f, xs), oracle_eval(po, f, xs)) < TOL


def test_eval_edge_points(q, po, eng):
    L = 10
    s = q.continuous_siteinds(q.named_grid((L, 1)))
    f = q.exp_itn(s, k=1.7, a=-0.3, c=0.9)
    h = 2.0 ** -L
    xs = np.array([[0.0], [h], [0.5], [1 - h], [np.nextafter(1.0, 0)], [1.0], [5.0], [h * 0.999]])
    y = q.evaluate(f, xs)
    xg = np.minimum(grid_of(xs[:, 0], L), 1 - h)
    assert np.allclose(y, 0.9 * np.exp(1.7 * xg - 0.3), rtol=1e-13)
    assert q.evaluate(f, 0.625) == pytest.approx(0.9 * math.exp(1.7 * 0.625 - 0.3), rel=1e-13)
    assert q.evaluate(f, np.zeros((0, 1))).shape == (0,)


def test_eval_invalid_coordinate(q, eng):
    s = q.continuous_siteinds(q.named_grid((8, 1)))
    f = q.exp_itn(s)
    with pytest.raises(q.DigitMapError):
        q.evaluate(f, np.array([[0.5], [-0.25]]))
    with pytest.raises(q.DigitMapError):
        q.evaluate(f, np.array([[float("nan")]]))
    # the status is cleared: a valid call afterwards succeeds
    assert q.evaluate(f, 0.5) == pytest.approx(math.exp(0.5), rel=1e-13)


def test_eval_device_tensors_deterministic(q, po, eng, rng):
    import torch
    s = q.continuous_siteinds(q.named_grid((30, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=16, rng=rng, std=0.25)
    xs = rng.uniform(0, 1, (100000, 2))
    xd = torch.tensor(xs, device="cuda")
    y1 = q.evaluate(f, xd).cpu().numpy()
    y2 = q.evaluate(f, xd).cpu().numpy()
    assert np.array_equal(y1, y2)
    assert rel_l2(y1, oracle_eval(po, f, xs)) < TOL


@pytest.mark.parametrize("prec,tol", [(1, 1e-10), (2, 1e-5)])
def test_eval_float32_paths(q, po, eng, prec, tol):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.laplacian_2d(npts=20000)
    y = q.evaluate(w.f, w.xs, precision=prec)
    ref = oracle_eval(po, w.f, w.xs)
    err = rel_l2(y, ref)
    print(f"precision {prec}: rel L2 error vs float64 oracle {err:.3e}")
    assert err < tol


# ------------------------------------------------------------------ integrate / inner
def test_integrate_chain_and_tree(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.sine_1d(npts=10)
    t = w.f.topo
    for take_sum, wt in ((False, 0.5), (True, 1.0)):
        got = q.integrate(w.f, take_sum=take_sum)
        ref = po.integrate(t.parent, t.sdim, w.f.cores, wt)
        assert got == pytest.approx(ref, rel=TOL)
    s = q.continuous_siteinds(q.named_comb_tree((3, 10)), map_dimension=2)
    f = q.exp_itn(s, dim=1) * q.exp_itn(s, dim=2)
    assert abs(q.integrate(f) - (math.e - 1) ** 2) < 1e-4  # test/test_integration.jl:28-38
    g = q.random_itn(s, link_space=5, rng=rng)
    assert q.integrate(g) == pytest.approx(po.integrate(g.topo.parent, g.topo.sdim, g.cores, 0.5), rel=TOL)


def test_inner_and_integrate_operator(q, po, eng, rng):
    s = q.continuous_siteinds(q.named_comb_tree((2, 6)))
    f = q.random_itn(s, link_space=4, rng=rng)
    g = q.random_itn(s, link_space=3, rng=rng)
    t = f.topo
    assert q.inner(f, g) == pytest.approx(po.inner(t.parent, t.sdim, f.cores, g.cores), rel=TOL)
    op = q.second_derivative_operator(s)
    C = po.contract(t.parent, t.sdim, op.cores, f.cores)
    ref = po.integrate(t.parent, t.sdim, C, 0.5)
    assert q.integrate(op, f) == pytest.approx(ref, rel=1e-11)


# ------------------------------------------------------------------ operate / truncate
def _compare_networks(q, po, g_dev, ref_cores, f, rng, tol=TOL, npts=4000):
    """Largest difference relative to the largest value, over random points plus
    points on the grid's edge lines (where boundary layers of operator results
    live).  A truncated operator result can be ~0 at interior points (only
    truncation residue), so a per-point relative error there is meaningless;
    this metric is relative to the function's scale."""
    nd = f.dimension()
    xs = rng.uniform(0, 1, (npts, nd))
    edge = []
    for d in range(nd):
        h = 2.0 ** -int(f.topo.Lc[d])
        for val in (0.0, h, 1 - 2 * h, 1 - h):
            e = rng.uniform(0, 1, (64, nd))
            e[:, d] = val
            edge.append(e)
    xs = np.concatenate([xs] + edge)
    y = q.evaluate(g_dev, xs)
    t = f.topo
    yr, _ = po.evaluate(t.parent, t.sdim, ref_cores, t.vcoord, t.vdigit, t.Lc, xs)
    return float(np.max(np.abs(y - yr)) / np.max(np.abs(yr)))


def test_apply_exact_matches_oracle(q, po, eng, rng):
    s = q.continuous_siteinds(q.named_comb_tree((2, 5)))
    f = q.random_itn(s, link_space=3, rng=rng)
    op = q.first_derivative_operator(s, left_boundary="Periodic", right_boundary="Neumann")
    g = q.operate(op, f, truncate=False)
    C, _ = _oracle_apply_truncate(po, f, op, -1, 0)
    assert _compare_networks(q, po, g, C, f, rng) < TOL
    assert list(g.bonds()) == [c.shape[0] for c in C]


def test_operate_truncate_sine_1d(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.sine_1d(npts=10)
    g, err = q.operate(w.op, w.f, cutoff=w.cutoff, return_error=True)
    ref, rerr = _oracle_apply_truncate(po, w.f, w.op, w.cutoff, 0)
    assert list(g.bonds()) == [c.shape[0] for c in ref]
    assert _compare_networks(q, po, g, ref, w.f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-6, abs=1e-20)
    # telescoping identity of the Dirichlet second derivative: exact for the
    # untruncated product, to the truncation error (cutoff 1e-12) after it
    h = 2.0 ** -12
    fx = workloads.sine_1d_function
    ident = -(fx(1 - h) + fx(0.0)) / h
    assert q.integrate(q.operate(w.op, w.f, truncate=False)) == pytest.approx(ident, rel=1e-11)
    assert q.integrate(g) == pytest.approx(ident, rel=1e-6)


def test_operate_truncate_laplacian_2d(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.laplacian_2d(npts=10)
    fc = q.truncate(w.f, cutoff=1e-16, maxdim=64)
    t = w.f.topo
    fref, _ = po.truncate(t.parent, t.sdim, w.f.cores, 1e-16, 64)
    assert list(fc.bonds()) == [c.shape[0] for c in fref]
    assert _compare_networks(q, po, fc, fref, w.f, rng) < TOL
    g = q.operate(w.op, fc, cutoff=w.cutoff, maxdim=w.maxdim)
    C = po.contract(t.parent, t.sdim, w.op.cores, fref)
    gref, _ = po.truncate(t.parent, t.sdim, C, w.cutoff, w.maxdim)
    assert list(g.bonds()) == [c.shape[0] for c in gref]
    assert _compare_networks(q, po, g, gref, w.f, rng) < 1e-11


@pytest.mark.parametrize("graph", ["chain", "comb"])
def test_truncate_random(q, po, eng, rng, graph):
    g = q.named_grid((14, 1)) if graph == "chain" else q.named_comb_tree((3, 5))
    s = q.continuous_siteinds(g)
    f = q.random_itn(s, link_space=12, rng=rng)
    tr, err = q.truncate(f, cutoff=1e-4, maxdim=6, return_error=True)
    ref, rerr = po.truncate(f.topo.parent, f.topo.sdim, f.cores, 1e-4, 6)
    assert list(tr.bonds()) == [c.shape[0] for c in ref]
    assert _compare_networks(q, po, tr, ref, f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-9)


@pytest.mark.parametrize("graph,maxdim", [("chain", None), ("comb", None), ("comb", 8)])
def test_truncate_keep_all_moves(q, po, eng, rng, graph, maxdim):
    # cutoff 0: the tall and square SVD moves whose rank the bound does not
    # cut keep Q R without a Jacobi (cores in LDS here); with maxdim 8 the
    # moves of rank above 8 still take the SVD.  Against the oracle's SVD.
    g = q.named_grid((14, 1)) if graph == "chain" else q.named_comb_tree((3, 5))
    s = q.continuous_siteinds(g)
    f = q.random_itn(s, link_space=12, rng=rng)
    tr, err = q.truncate(f, cutoff=0.0, maxdim=maxdim, return_error=True)
    ref, rerr = po.truncate(f.topo.parent, f.topo.sdim, f.cores, 0.0, maxdim or (1 << 30))
    assert list(tr.bonds()) == [c.shape[0] for c in ref]
    assert _compare_networks(q, po, tr, ref, f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-9, abs=1e-24)


@pytest.mark.parametrize("bond,cutoff", [(130, 1e-10), (100, 1e-3), (130, 0.0)])
def test_truncate_large_matrices(q, po, eng, rng, bond, cutoff):
    # moves whose matrix does not fit LDS (e.g. 260 x 130): blocked QR with panels in LDS,
    # U of the tall SVD moves formed the same way.  Cutoff 0 with no maxdim: the
    # tall SVD moves can drop nothing, so they keep Q R (no Jacobi); the oracle
    # runs the SVD, and the bonds, values and error must still agree
    s = q.continuous_siteinds(q.named_grid((12, 1)))
    f = q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond))
    tr, err = q.truncate(f, cutoff=cutoff, return_error=True)
    ref, rerr = po.truncate(f.topo.parent, f.topo.sdim, f.cores, cutoff)
    assert list(tr.bonds()) == [c.shape[0] for c in ref]
    assert _compare_networks(q, po, tr, ref, f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-8, abs=1e-20)


def test_truncate_preconditioned_wide_moves(q, po, eng, rng):
    # an 18-site chain of bond 256 truncated to maxdim 100: cores beyond LDS,
    # so the launch carries helper workgroups, and the SVD moves back from the
    # last site meet 200 x 256 matrices (wide, >= 192 rows), whose Jacobi runs
    # QR-preconditioned (X^T = Q1 R1, R1^T = Q2 R2, rounds on R2); the QR moves
    # keep their reflectors for the SVD moves to apply
    s = q.continuous_siteinds(q.named_grid((18, 1)))
    f = q.random_itn(s, link_space=256, rng=rng, std=1.0 / math.sqrt(256))
    tr, err = q.truncate(f, cutoff=0.0, maxdim=100, return_error=True)
    ref, rerr = po.truncate(f.topo.parent, f.topo.sdim, f.cores, 0.0, 100)
    assert list(tr.bonds()) == [c.shape[0] for c in ref]
    assert max(tr.bonds()) == 100
    assert _compare_networks(q, po, tr, ref, f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-9)


# ------------------------------------------------------------------ comb tree (config-3 shape, reduced)
def test_comb_operate_and_integrate(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.comb_3d(bond=8, digits=6)
    f = w.f
    t = f.topo
    assert q.integrate(f) == pytest.approx(po.integrate(t.parent, t.sdim, f.cores, 0.5), rel=TOL)
    g = q.operate(w.op, f, cutoff=w.cutoff, maxdim=w.maxdim)
    C = po.contract(t.parent, t.sdim, w.op.cores, f.cores)
    gref, _ = po.truncate(t.parent, t.sdim, C, w.cutoff, w.maxdim)
    assert list(g.bonds()) == [c.shape[0] for c in gref]
    assert _compare_networks(q, po, g, gref, f, rng) < 1e-11
    assert q.integrate(g) == pytest.approx(po.integrate(t.parent, t.sdim, gref, 0.5), rel=1e-11)


def _star_network(q, rng):
    """A vertex x with three arms of six digits: arm a ends at the root, x's
    bonds are 24 (to a), 20 and 25, so moves at x have 1000-1200 rows: too tall
    for the two-copy LDS panel."""
    verts = [("a", j) for j in range(6, 0, -1)] + ["x"] + [(arm, j) for arm in "bc" for j in range(1, 7)]
    edges = ([(("a", j), ("a", j - 1)) for j in range(6, 1, -1)] + [(("a", 1), "x"), ("x", ("b", 1)), ("x", ("c", 1))]
             + [((arm, j), (arm, j + 1)) for arm in "bc" for j in range(1, 6)])
    s = q.continuous_siteinds(q.NamedGraph(verts, edges), [verts])
    topo = q.Topology(s)
    to_parent = {("a", 6): 1, ("a", 5): 2, ("a", 4): 4, ("a", 3): 8, ("a", 2): 16, ("a", 1): 24, "x": 24}
    for arm, top in (("b", 20), ("c", 25)):
        for j in range(1, 7):
            to_parent[(arm, j)] = min(top, 2 ** (7 - j))
    bond = np.array([to_parent[v] for v in topo.names], dtype=np.int64)
    cores = [rng.standard_normal(topo.core_shape(v, bond)) / math.sqrt(24) for v in range(topo.nv)]
    return q.ITensorNetworkFunction(s, cores, topo)


@pytest.mark.parametrize("case", ["star", "comb"])
def test_truncate_global_workspace(q, po, eng, rng, case):
    # moves whose matrices do not fit LDS: at the star's centre (1000-1200 rows: one LDS
    # panel copy, or panels in global memory) and at the comb's backbone (3200 x 40)
    if case == "star":
        f = _star_network(q, rng)
        g, err = q.truncate(f, cutoff=1e-10, return_error=True)
        ref, rerr = po.truncate(f.topo.parent, f.topo.sdim, f.cores, 1e-10)
    else:
        from itensornumericalanalysis_jl_amd import workloads
        w = workloads.comb_3d(bond=40)
        f = w.f
        t = f.topo
        g, err = q.operate(w.op, f, cutoff=w.cutoff, maxdim=w.maxdim, return_error=True)
        C = po.contract(t.parent, t.sdim, w.op.cores, f.cores)
        ref, rerr = po.truncate(t.parent, t.sdim, C, w.cutoff, w.maxdim)
    assert list(g.bonds()) == [c.shape[0] for c in ref]
    assert _compare_networks(q, po, g, ref, f, rng) < 1e-11
    assert err == pytest.approx(rerr, rel=1e-8, abs=1e-20)


@pytest.mark.parametrize("case", ["comb", "star", "no_digit"])
def test_integrate_dataflow(q, po, eng, rng, case):
    # networks of >= 2^15 core elements take the many-workgroup dataflow kernel
    if case == "comb":
        s = q.continuous_siteinds(q.named_comb_tree((3, 6)))
        bond = 24
    elif case == "star":  # a digit-less centre with four arms: a vertex with four children
        verts = ["c"] + [(a, j) for a in range(1, 5) for j in range(1, 6)]
        edges = [("c", (a, 1)) for a in range(1, 5)] + [((a, j), (a, j + 1)) for a in range(1, 5)
                                                         for j in range(1, 5)]
        s = q.continuous_siteinds(q.NamedGraph(verts, edges), [[(a, j) for j in range(1, 6)] for a in range(1, 5)])
        bond = 14
    else:  # the third tooth carries no digits
        s = q.continuous_siteinds(q.named_comb_tree((3, 6)), [[(1, j) for j in range(1, 7)],
                                                              [(2, j) for j in range(1, 7)]])
        bond = 24
    f = q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond))
    assert sum(c.size for c in f.cores) >= 1 << 15
    t = f.topo
    for take_sum, wgt in ((False, 0.5), (True, 1.0)):
        ref = po.integrate(t.parent, t.sdim, f.cores, wgt)
        assert q.integrate(f, take_sum=take_sum) == pytest.approx(ref, rel=TOL)


def test_comb_full_size_integrate(q, po, eng):
    # 3 x 16 digits, bond 128: 45 MB of cores
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.comb_3d()
    t = w.f.topo
    ref = po.integrate(t.parent, t.sdim, w.f.cores, 0.5)
    for _ in range(3):  # repeated launches: the counters are left zero by each launch
        assert q.integrate(w.f) == pytest.approx(ref, rel=TOL)


def test_comb_full_size_operate_error_bound(q, eng):
    # The 3D comb at full size: 3 x 16 digits, bond 128, so W f has bond 384.
    # The oracle needs about 75 s for this operate.  So the check uses the
    # truncation's own identities instead, from inner products on the GPU:
    # - ||W f - g||^2 is at most the sum of the discarded weights times
    #   ||W f||^2, and close to it (the sequential bound is tight here);
    # - g is nearly the orthogonal projection of W f: <W f, g> = <g, g>.
    from itensornumericalanalysis_jl_amd import workloads
    w = workloads.comb_3d()
    Wf = q.operate(w.op, w.f, truncate=False)
    g, err = q.operate(w.op, w.f, cutoff=w.cutoff, maxdim=w.maxdim, return_error=True)
    assert max(Wf.bonds()) == 384 and max(g.bonds()) == 128
    nw, cr, ng = q.inner(Wf, Wf), q.inner(Wf, g), q.inner(g, g)
    d2 = (nw - 2 * cr + ng) / nw
    assert 0 < d2 <= err * (1 + 1e-9)
    assert d2 == pytest.approx(err, rel=1e-3)
    assert cr == pytest.approx(ng, rel=1e-9)


# ------------------------------------------------------------------ batches (config-5 shape, reduced)
def test_operate_batch_and_integrate_batch(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import workloads
    ws = workloads.fredholm_batch(nfun=12, npts=64, seed=5)
    for w in ws:  # keep the oracle fast: cap the bond
        if w.f.maxlinkdim() > 24:
            w.f = q.random_itn(w.s, link_space=24, rng=rng)
    gs, errs = q.operate_batch([w.op for w in ws], [w.f for w in ws], cutoff=1e-10)
    ints = q.integrate_batch(gs)
    for w, g, e, integ in zip(ws, gs, errs, ints):
        t = w.f.topo
        C = po.contract(t.parent, t.sdim, w.op.cores, w.f.cores)
        gref, rerr = po.truncate(t.parent, t.sdim, C, 1e-10)
        assert list(g.bonds()) == [c.shape[0] for c in gref]
        assert _compare_networks(q, po, g, gref, w.f, rng, npts=512) < 1e-11
        assert integ == pytest.approx(po.integrate(t.parent, t.sdim, gref, 0.5), rel=1e-11)
        assert e == pytest.approx(rerr, rel=1e-6, abs=1e-18)


def test_operate_batch_order_does_not_matter(q, eng, rng):
    # the batch truncation dispatches its costliest networks first; each network
    # is swept by one workgroup on its own, so the order of the list changes no bit
    from itensornumericalanalysis_jl_amd import workloads
    ws = workloads.fredholm_batch(nfun=10, npts=16, seed=6)
    ops, fs = [w.op for w in ws], [w.f for w in ws]
    ga, ea = q.operate_batch(ops, fs, cutoff=1e-10)
    gb, eb = q.operate_batch(ops[::-1], fs[::-1], cutoff=1e-10)
    for x, y, e1, e2 in zip(ga, gb[::-1], ea, eb[::-1]):
        assert list(x.bonds()) == list(y.bonds())
        assert e1 == e2
        for c1, c2 in zip(x.cores, y.cores):
            assert np.array_equal(c1, c2)


def test_tune_evaluate(q, po, eng, rng):
    # the autotuner times the register kernel's launch shapes and later calls use the fastest;
    # every shape runs the same arithmetic per point, so the values do not change
    import torch
    s = q.continuous_siteinds(q.named_grid((16, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=12, rng=rng, std=1.0 / math.sqrt(12))
    xs = rng.uniform(0, 1, (200003, 2))
    xd = torch.tensor(xs, device="cuda")
    before = q.evaluate(f, xd).cpu().numpy()
    cfg = q.tune_evaluate(f, xd)
    assert cfg["points_per_lane"] in (1, 2) and cfg["workgroup"] in (64, 128, 256) and cfg["ms"] > 0
    assert cfg["table_digits"] in (0, 8, 10, 12, 14, 16)  # at most half of the 32 digit sites
    after = q.evaluate(f, xd).cpu().numpy()
    # another table size changes the rounding order, not the values
    assert np.allclose(before, after, rtol=1e-13, atol=1e-13 * np.abs(before).max())
    assert rel_l2(after[:5000], oracle_eval(po, f, xs[:5000])) < TOL
    # a tree does not take the register kernel: nothing to tune
    st = q.continuous_siteinds(q.named_comb_tree((2, 4)))
    ft = q.random_itn(st, link_space=3, rng=rng)
    xt = torch.tensor(rng.uniform(0, 1, (100, ft.dimension())), device="cuda")
    assert q.tune_evaluate(ft, xt)["workgroup"] == 0


def test_evaluate_batch(q, po, eng, rng):
    # uneven digit counts and bonds (all four register classes, one LDS-kernel network), ragged point counts
    import torch
    fs, xs = [], []
    for i, (L, bond) in enumerate([(8, 3), (12, 7), (20, 16), (16, 29), (10, 40), (14, 1), (9, 12)]):
        s = q.continuous_siteinds(q.named_grid((L, 1)))
        fs.append(q.random_itn(s, link_space=bond, rng=rng, std=1.0 / math.sqrt(bond)))
        xs.append(rng.uniform(0, 1, (100 + 777 * i, 1)))
    outs = q.evaluate_batch(fs, [torch.tensor(x, device="cuda") for x in xs])
    eng.sync()
    for f, x, y in zip(fs, xs, outs):
        yb = y.cpu().numpy()
        assert rel_l2(yb, oracle_eval(po, f, x)) < TOL
        assert np.array_equal(yb, q.evaluate(f, x))  # same arithmetic as the single-network launch


# ------------------------------------------------------------------ gradient
@pytest.mark.parametrize("bond", [3, 8, 16, 32, 40])
def test_gradient_matches_oracle(q, po, eng, rng, bond):
    # bonds <= 32 run the register kernel (8/16/32 lanes of state), 40 the general one
    s = q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=bond, rng=rng, std=0.4)
    xs = rng.uniform(0, 1, (5000, 2))
    xs[:3] = [[0.0, 0.0], [1.0, 1.0], [0.5, 0.999999]]
    yb = rng.standard_normal(5000)
    g = q.eval_grad(f, xs, yb)
    t = f.topo
    ref = po.eval_grad(t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc, xs, yb)
    num = math.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(g, ref)))
    den = math.sqrt(sum(np.sum(b ** 2) for b in ref))
    assert num / den < TOL
    g2 = q.eval_grad(f, xs, yb)
    assert all(np.array_equal(a, b) for a, b in zip(g, g2))  # deterministic reduction

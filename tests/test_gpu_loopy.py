"""Belief-propagation evaluation on layouts with loops (qtt_bp_eval_host)
against the oracle's BP (same schedule and stopping rule), exact cases
(trees, bond 1), the single-loop eigenvalue and the reference's constant on a
3 x 3 grid (test/test_realitensorfunction.jl:41-50)."""
import numpy as np
import pytest

from helpers import oracle_eval

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _args(f):
    return len(f.names), f.edges, f.sdim, f.vcoord, f.vdigit, f.Lc, f.cores


def _positive(q, s, bond, rng):
    f = q.random_itn(s, link_space=bond, rng=rng)
    f.cores = [np.abs(c) + 0.2 for c in f.cores]
    return f


@pytest.mark.parametrize("case", ["grid3_1d", "grid4_2d", "grid3_bond4"])
def test_bp_matches_oracle(q, po, eng, rng, case):
    if case == "grid3_1d":
        s, bond = q.continuous_siteinds(q.named_grid((3, 3))), 2
    elif case == "grid4_2d":
        s, bond = q.continuous_siteinds(q.named_grid((4, 4)), map_dimension=2), 3
    else:
        s, bond = q.continuous_siteinds(q.named_grid((3, 3))), 4
    f = _positive(q, s, bond, rng)
    xs = rng.uniform(0, 1, (300, s.dimension()))
    y, sweeps = q.evaluate_bp(f, xs, return_sweeps=True)
    ref, rsweeps = po.bp_evaluate(*_args(f), xs[:30])
    assert np.allclose(y[:30], ref, rtol=1e-10), case
    assert abs(sweeps - rsweeps) <= 2 and sweeps < 200
    # BP is an estimate on loops: it differs from the exact scalar, but not wildly here
    ex = po.loopy_exact(*_args(f), xs[:20])
    assert np.all(np.abs(y[:20] / ex - 1) < 0.5)
    assert np.allclose(q.evaluate(f, xs), y, rtol=0, atol=0)  # evaluate dispatches to BP


def test_bp_exact_cases(q, po, eng, rng):
    from itensornumericalanalysis_jl_amd import loopy
    # a tree: BP is exact
    s = q.continuous_siteinds(q.named_comb_tree((3, 4)), map_dimension=2)
    t = q.random_itn(s, link_space=3, rng=rng)
    xs = rng.uniform(0, 1, (500, 2))
    y = q.evaluate_bp(loopy.from_tree(t), xs)
    assert np.allclose(y, oracle_eval(po, t, xs), rtol=1e-11, atol=1e-12)
    # bond 1 on a grid: exact
    s2 = q.continuous_siteinds(q.named_grid((4, 4)))
    f = q.random_itn(s2, link_space=1, rng=rng)
    xs1 = rng.uniform(0, 1, (200, 1))
    assert np.allclose(q.evaluate_bp(f, xs1), po.loopy_exact(*_args(f), xs1), rtol=1e-12)
    # one loop: the dominant eigenvalue of the ring product
    verts = [1, 2, 3, 4, 5]
    g = q.NamedGraph(verts, [(1, 2), (2, 3), (3, 4), (4, 5), (5, 1)])
    sr = q.IndsNetworkMap(g, [verts])
    cores = [np.abs(rng.standard_normal((2, 4, 4))) + 0.1 for _ in verts]
    fr = q.LoopyNetworkFunction(sr, cores, [4] * 5)
    x = 0.40625
    d = [po.digit(x, 5, j) for j in range(1, 6)]
    T = [c[k] for c, k in zip(cores, d)]
    # legs: v1 (e0, e4), v2 (e0, e1), v3 (e1, e2), v4 (e2, e3), v5 (e3, e4)
    P = T[0].T @ T[1] @ T[2] @ T[3] @ T[4]
    assert q.evaluate_bp(fr, x) == pytest.approx(np.max(np.linalg.eigvals(P).real), rel=1e-10)
    # the reference's constant on a 3 x 3 grid
    sc = q.continuous_siteinds(q.named_grid((3, 3)))
    assert q.evaluate(q.const_itn(sc, c=1.5), 0.5) == pytest.approx(1.5, rel=1e-14)


def test_bp_invalid_point(q, eng, rng):
    s = q.continuous_siteinds(q.named_grid((3, 3)))
    f = _positive(q, s, 2, rng)
    with pytest.raises(q.DigitMapError):
        q.evaluate_bp(f, np.array([[0.5], [np.nan]]))
    assert np.isfinite(q.evaluate_bp(f, 0.5))

"""The oracle's belief propagation (oracle/pyoracle.py bp_evaluate), pinned
before it checks the device: exact on trees (against the C oracle's
evaluate), exact on bond-1 networks with loops (against brute force), the
dominant eigenvalue of the ring product on a single loop (the known fixed
point of BP on one cycle, where the exact value is the trace), and the
reference's constant on a 3 x 3 grid (test/test_realitensorfunction.jl:41-50)."""
import numpy as np
import pytest

from helpers import oracle_eval


def _args(f):
    return len(f.names), f.edges, f.sdim, f.vcoord, f.vdigit, f.Lc, f.cores


def test_bp_exact_on_trees(q, po, rng):
    from itensornumericalanalysis_jl_amd import loopy
    s = q.continuous_siteinds(q.named_comb_tree((3, 3)), map_dimension=2)
    t = q.random_itn(s, link_space=3, rng=rng)
    f = loopy.from_tree(t)
    xs = rng.uniform(0, 1, (40, 2))
    y, sweeps = po.bp_evaluate(*_args(f), xs)
    assert np.allclose(y, oracle_eval(po, t, xs), rtol=1e-11, atol=1e-12)
    assert np.allclose(po.loopy_exact(*_args(f), xs), oracle_eval(po, t, xs), rtol=1e-12, atol=1e-13)


def test_bp_exact_for_bond_one_loops(q, po, rng):
    s = q.continuous_siteinds(q.named_grid((3, 3)))
    f = q.random_itn(s, link_space=1, rng=rng)
    xs = rng.uniform(0, 1, (30, 1))
    y, _ = po.bp_evaluate(*_args(f), xs)
    assert np.allclose(y, po.loopy_exact(*_args(f), xs), rtol=1e-12)


def test_bp_single_loop_gives_dominant_eigenvalue(q, po, rng):
    verts = [1, 2, 3, 4]
    g = q.NamedGraph(verts, [(1, 2), (2, 3), (3, 4), (4, 1)])
    s = q.IndsNetworkMap(g, [verts])
    from itensornumericalanalysis_jl_amd.loopy import LoopyNetworkFunction
    cores = [np.abs(rng.standard_normal((2, 3, 3))) + 0.1 for _ in verts]
    f = LoopyNetworkFunction(s, cores, [3, 3, 3, 3])
    x = np.array([[0.3]])
    T = [c[d] for c, d in zip(cores, [po.digit(0.3, 4, j) for j in (1, 2, 3, 4)])]
    # legs: v0 (e0, e3), v1 (e0, e1), v2 (e1, e2), v3 (e2, e3): product around the ring
    P = T[0].T @ T[1] @ T[2] @ T[3]
    lam = np.linalg.eigvals(P)
    y, sweeps = po.bp_evaluate(*_args(f), x)
    assert y[0] == pytest.approx(np.max(lam.real), rel=1e-10)
    assert po.loopy_exact(*_args(f), x)[0] == pytest.approx(np.trace(P), rel=1e-12)


def test_bp_reference_constant_grid(q, po):
    s = q.continuous_siteinds(q.named_grid((3, 3)))
    f = q.const_itn(s, c=1.5)
    y, _ = po.bp_evaluate(*_args(f), np.array([[0.5]]))
    assert y[0] == pytest.approx(1.5, rel=1e-14)
    with pytest.raises(NotImplementedError):
        q.const_itn(s, c=1.5, linkdim=4)

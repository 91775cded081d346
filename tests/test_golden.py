"""Checks against the committed fixtures of tests/golden (written by
tests/golden/make_fixtures.py from definitions and closed forms).

CPU: the host operator builders and the oracle reproduce the fixtures.
GPU (marked): the engine reproduces them through the C-ABI.
"""
import os

import numpy as np
import pytest

from oracle import dense

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cores(fx):
    n = sum(1 for k in fx if k.startswith("core"))
    return [fx[f"core{v}"] for v in range(n)]


def _digit_order_matrix(op, L):
    t = op.topo
    M = dense.dense_operator(t.parent, op.cores)
    order = np.argsort(t.vdigit)
    M = M.reshape((2,) * (2 * L))
    return np.transpose(M, list(order) + [L + o for o in order]).reshape(2 ** L, 2 ** L)


# ---------------------------------------------------------------- CPU
def test_fixture_shift_operators(q):
    fx = _load("shift_ops_L6.npz")
    s = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    for boundary in ("Dirichlet", "Periodic", "Neumann"):
        for n in (0, 1):
            for fwd, build in (("forward", q.forward_shift_op), ("backward", q.backward_shift_op)):
                M = _digit_order_matrix(build(s, boundary=boundary, n=n), 6)
                assert np.allclose(M, fx[f"{fwd}_{boundary}_n{n}"], atol=1e-12), (fwd, boundary, n)
    sc = q.continuous_siteinds(q.named_grid((6, 1)))
    for p, name in enumerate(["first", "second", "third", "fourth"], start=1):
        M = _digit_order_matrix(getattr(q, f"{name}_derivative_operator")(sc), 6)
        ref = fx[f"derivative_{p}"]
        assert np.allclose(M, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), name


def _comb_apply_oracle(po, fx, op):
    C = po.contract(fx["parent"], fx["sdim"], op.cores, _cores(fx))
    T = dense.dense_tensor(fx["parent"], C)
    return np.transpose(T, np.argsort(fx["vdigit"])).reshape(-1)


def test_fixture_comb_shifts_oracle(q, po):
    # test/test_operators.jl:168-282 on all 64 grid points
    fx = _load("comb_shifts_L6.npz")
    s = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    T = dense.dense_tensor(fx["parent"], _cores(fx))
    assert np.allclose(np.transpose(T, np.argsort(fx["vdigit"])).reshape(-1), fx["psi"], rtol=1e-14)
    for boundary in ("Dirichlet", "Periodic", "Neumann"):
        for n in (0, 1):
            for fwd, build in (("forward", q.forward_shift_op), ("backward", q.backward_shift_op)):
                got = _comb_apply_oracle(po, fx, build(s, boundary=boundary, n=n))
                assert np.allclose(got, fx[f"{fwd}_{boundary}_n{n}"], atol=1e-13), (fwd, boundary, n)


def test_fixture_c1_oracle(q, po):
    fx = _load("c1_sine.npz")
    cores = _cores(fx)
    y, rc = po.evaluate(fx["parent"], fx["sdim"], cores, fx["vcoord"], fx["vdigit"], fx["Lc"], fx["xs"])
    assert rc == 0
    assert np.allclose(y, fx["f"], rtol=1e-13, atol=1e-13)
    assert po.integrate(fx["parent"], fx["sdim"], cores, 0.5) == pytest.approx(float(fx["integral_f"]), rel=1e-12)
    s = q.continuous_siteinds(q.named_grid((12, 1)))
    op = q.second_derivative_operator(s)
    C = po.contract(fx["parent"], fx["sdim"], op.cores, cores)
    yd, _ = po.evaluate(fx["parent"], fx["sdim"], C, fx["vcoord"], fx["vdigit"], fx["Lc"], fx["xs"])
    assert np.max(np.abs(yd - fx["d2f"])) <= 1e-9 * np.max(np.abs(fx["d2f"]))
    assert po.integrate(fx["parent"], fx["sdim"], C, 0.5) == pytest.approx(float(fx["integral_d2f"]), rel=1e-11)


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _fixture_function(q, fx, s):
    topo = q.Topology(s)
    assert np.array_equal(topo.parent, fx["parent"]) and np.array_equal(topo.vdigit, fx["vdigit"])
    return q.ITensorNetworkFunction(s, _cores(fx), topo)


@pytest.mark.gpu
def test_fixture_comb_shifts_engine(q, eng):
    fx = _load("comb_shifts_L6.npz")
    s = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    psi = _fixture_function(q, fx, s)
    grid = fx["grid"][:, None]
    assert np.allclose(q.evaluate(psi, grid), fx["psi"], rtol=1e-14)
    for boundary in ("Dirichlet", "Periodic", "Neumann"):
        for n in (0, 1):
            for fwd, build in (("forward", q.forward_shift_op), ("backward", q.backward_shift_op)):
                g = q.operate(build(s, boundary=boundary, n=n), psi, cutoff=1e-12)
                got = q.evaluate(g, grid)
                assert np.allclose(got, fx[f"{fwd}_{boundary}_n{n}"], atol=1e-12), (fwd, boundary, n)


@pytest.mark.gpu
def test_fixture_c1_engine(q, eng):
    fx = _load("c1_sine.npz")
    s = q.continuous_siteinds(q.named_grid((12, 1)))
    f = _fixture_function(q, fx, s)
    assert np.allclose(q.evaluate(f, fx["xs"]), fx["f"], rtol=1e-13, atol=1e-13)
    assert q.integrate(f) == pytest.approx(float(fx["integral_f"]), rel=1e-12)
    op = q.second_derivative_operator(s)
    exact = q.operate(op, f, truncate=False)
    yd = q.evaluate(exact, fx["xs"])
    assert np.max(np.abs(yd - fx["d2f"])) <= 1e-9 * np.max(np.abs(fx["d2f"]))
    assert q.integrate(exact) == pytest.approx(float(fx["integral_d2f"]), rel=1e-11)
    g = q.operate(op, f, cutoff=1e-12)  # the reference's cutoff: 1e-6 of the norm
    assert np.max(np.abs(q.evaluate(g, fx["xs"]) - fx["d2f"])) <= 1e-5 * np.max(np.abs(fx["d2f"]))

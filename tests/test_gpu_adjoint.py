"""The adjoint kernels against the oracle: the evaluate gradient on general
trees (float64), the float32 chain gradient (error against the float64
oracle stated), and the pullback of the exact operator application."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _rel(a_list, b_list):
    num = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(a_list, b_list)))
    den = math.sqrt(sum(float(np.sum(b ** 2)) for b in b_list))
    return num / den


def _star(q):
    # a digit-less centre with four arms of five digits (a vertex with four children)
    verts = ["c"] + [(a, j) for a in range(1, 5) for j in range(1, 6)]
    edges = [("c", (a, 1)) for a in range(1, 5)] + [((a, j), (a, j + 1)) for a in range(1, 5) for j in range(1, 5)]
    return q.continuous_siteinds(q.NamedGraph(verts, edges), [[(a, j) for j in range(1, 6)] for a in range(1, 5)])


@pytest.mark.parametrize("case", ["comb", "comb_2d", "star"])
def test_gradient_tree_matches_oracle(q, po, eng, rng, case):
    if case == "comb":
        s = q.continuous_siteinds(q.named_comb_tree((3, 5)))
        bond = 4
    elif case == "comb_2d":
        s = q.continuous_siteinds(q.named_comb_tree((4, 4)), map_dimension=2)
        bond = 3
    else:
        s = _star(q)
        bond = 3
    f = q.random_itn(s, link_space=bond, rng=rng, std=0.5)
    assert not f.topo.is_chain()
    nd = f.dimension()
    xs = rng.uniform(0, 1, (3000, nd))
    xs[0] = 0.0
    xs[1] = 1.0
    yb = rng.standard_normal(3000)
    g = q.eval_grad(f, xs, yb)
    t = f.topo
    ref = po.eval_grad(t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc, xs, yb)
    assert _rel(g, ref) < TOL
    g2 = q.eval_grad(f, xs, yb)
    assert all(np.array_equal(a, b) for a, b in zip(g, g2))  # fixed-order reduction


def test_gradient_float32_fit_2d(q, po, eng):
    # the 2D fit (2 x 14 digits, bond 32): float32 gradient against the float64
    # oracle on a subset, and against the float64 engine on all 2^20 points
    import ctypes
    import torch
    from itensornumericalanalysis_jl_amd import _capi, workloads
    w = workloads.fit_2d()
    f = w.f
    n = len(w.xs)
    xs = torch.tensor(w.xs, device="cuda")
    y = q.evaluate(f, xs)
    ybar = (2.0 / n) * (y - torch.tensor(w.extra["target"], device="cuda"))
    ns = 1 << 14
    g32 = q.eval_grad(f, w.xs[:ns], ybar[:ns].cpu().numpy(), precision=q.Precision.F32)
    t = f.topo
    ref = po.eval_grad(t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc, w.xs[:ns], ybar[:ns].cpu().numpy())
    e_sub = _rel(g32, ref)
    dn = q.api.device_network(f)

    def full(prec):
        out = torch.empty(dn.size(), dtype=torch.float64, device="cuda")
        eng.check(_capi.lib().qtt_eval_grad(eng.handle, dn.handle, ctypes.c_void_p(xs.data_ptr()),
                                            ctypes.c_void_p(ybar.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                            prec))
        eng.sync()
        return out
    G64, G32 = full(0), full(2)
    e_full = (torch.linalg.norm(G32 - G64) / torch.linalg.norm(G64)).item()
    print(f"float32 gradient: rel error {e_sub:.2e} (2^14 points vs oracle), {e_full:.2e} (2^20 points vs float64)")
    assert e_sub < 1e-5 and e_full < 1e-5
    assert e_full > 1e-10  # it is the float32 arithmetic that ran


@pytest.mark.parametrize("case", ["chain_d2", "comb_d1", "laplacian_2d"])
def test_apply_grad_matches_oracle(q, po, eng, rng, case):
    if case == "chain_d2":
        s = q.continuous_siteinds(q.named_grid((10, 1)))
        f = q.random_itn(s, link_space=5, rng=rng)
        op = q.second_derivative_operator(s)
    elif case == "comb_d1":
        s = q.continuous_siteinds(q.named_comb_tree((3, 4)))
        f = q.random_itn(s, link_space=3, rng=rng)
        op = q.first_derivative_operator(s, left_boundary="Periodic", right_boundary="Neumann")
    else:
        s = q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
        f = q.random_itn(s, link_space=4, rng=rng)
        op = q.laplacian_operator(s)
    t = f.topo
    C = po.contract(t.parent, t.sdim, op.cores, f.cores)
    G = [rng.standard_normal(c.shape) for c in C]
    fb, wb = q.apply_grad(op, f, G)
    rfb, rwb = po.contract_adjoint(t.parent, t.sdim, op.cores, f.cores, G)
    assert _rel(fb, rfb) < TOL
    assert _rel(wb, rwb) < TOL


def test_apply_grad_chain_rule(q, po, eng, rng):
    # the rrule of evaluate(operate(op, f; exact), xs): eval_grad on the product,
    # then the pullback of the product, against the same chain on the oracle
    s = q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=6, rng=rng, std=0.4)
    op = q.laplacian_operator(s)
    g = q.operate(op, f, truncate=False)
    xs = rng.uniform(0, 1, (2000, 2))
    yb = rng.standard_normal(2000)
    gb = q.eval_grad(g, xs, yb)
    fb, _ = q.apply_grad(op, f, gb, want_op=False)
    t = f.topo
    C = po.contract(t.parent, t.sdim, op.cores, f.cores)
    rg = po.eval_grad(t.parent, t.sdim, C, t.vcoord, t.vdigit, t.Lc, xs, yb)
    rfb, _ = po.contract_adjoint(t.parent, t.sdim, op.cores, f.cores, rg, want_w=False)
    assert _rel(fb, rfb) < 1e-11

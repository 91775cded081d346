"""Hadamard product, sum and partial integration on the device against the
oracle, and the Fredholm iteration of examples/fredholm_solver.jl:43-65 run
entirely on the device against the same iteration on the dense grid.

Bars: cores of products and sums are exact (one multiplication or a copy per
element, so bit-equal to the numpy oracle); values 1e-12 relative."""
import numpy as np
import pytest

from helpers import grid_of, oracle_eval

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _comb2(q, L):
    return q.continuous_siteinds(q.named_comb_tree((2, L)), [[(i, j) for j in range(1, L + 1)] for i in (1, 2)])


def _layout(q, case):
    if case == "chain":
        return q.continuous_siteinds(q.named_grid((11, 1)))
    if case == "chain2d":
        return q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
    return q.continuous_siteinds(q.named_comb_tree((3, 4)), [[(i, j) for j in range(1, 5)] for i in (1, 2, 3)])


def _host(q, f):
    """A host copy of a device network (its cores downloaded)."""
    return q.ITensorNetworkFunction(f.s, [c.copy() for c in f.cores], f.topo)


@pytest.mark.parametrize("case", ["chain", "chain2d", "comb"])
def test_multiply_and_add_same_layout(q, po, eng, rng, case):
    s = _layout(q, case)
    f = q.random_itn(s, link_space=5, rng=rng)
    g = q.random_itn(s, link_space=3, rng=rng)
    t = f.topo
    gf = q.multiply(g, f)
    H = po.hadamard(t.parent, t.sdim, g.cores, f.cores)
    for a, b in zip(gf.cores, H):
        assert np.array_equal(a, b)
    fs = f + g
    S = po.direct_sum(t.parent, t.sdim, f.cores, g.cores)
    for a, b in zip(fs.cores, S):
        assert np.array_equal(a, b)
    xs = rng.uniform(0.0, 1.0, (300, s.dimension()))
    yf, yg = oracle_eval(po, f, xs), oracle_eval(po, g, xs)
    assert np.allclose(q.evaluate(gf, xs), yf * yg, rtol=1e-12, atol=1e-12 * np.abs(yf * yg).max())
    assert np.allclose(q.evaluate(fs, xs), yf + yg, rtol=1e-12, atol=1e-12 * np.abs(yf + yg).max())
    # operands already on the device: a product of products
    ggf = q.multiply(g, gf)
    assert np.allclose(q.evaluate(ggf, xs), yg * yg * yf, rtol=1e-11, atol=1e-12 * np.abs(yg * yg * yf).max())


def test_multiply_parts_of_a_layout(q, po, eng, rng):
    """Functions on the two teeth of a comb (disjoint vertex sets) and on the
    whole comb: the product lives on the union, joined by the comb's edges."""
    s = _comb2(q, 5)
    s1, s2 = q.reduced_indsnetworkmap(s, 1), q.reduced_indsnetworkmap(s, 2)
    a = q.random_itn(s1, link_space=3, rng=rng)
    b = q.random_itn(s2, link_space=2, rng=rng)
    h = q.random_itn(s, link_space=2, rng=rng)
    ab = q.multiply(a, b)
    assert ab.s.same_layout(s)
    xs = rng.uniform(0.0, 1.0, (256, 2))
    ya = oracle_eval(po, a, xs[:, :1])
    yb = oracle_eval(po, b, xs)  # s2 keeps dimension 1 without digits: its value is ignored
    yh = oracle_eval(po, h, xs)
    assert np.allclose(q.evaluate(ab, xs), ya * yb, rtol=1e-12, atol=1e-12 * np.abs(ya * yb).max())
    bh = q.multiply(b, h)
    assert np.allclose(q.evaluate(bh, xs), yb * yh, rtol=1e-12, atol=1e-12 * np.abs(yb * yh).max())


def _brute_partial(po, f, dims, take_sum=False):
    """Values of the partial integral at every grid point of the remaining
    dimensions, by summing the oracle's values over the integrated grid."""
    s = f.s
    rest = [d for d in s.dimensions() if d not in dims]
    Ls = {d: s.ndigits(d) for d in s.dimensions()}
    rest_grid = np.stack(np.meshgrid(*[np.arange(2 ** Ls[d]) / 2.0 ** Ls[d] for d in rest], indexing="ij"),
                         -1).reshape(-1, len(rest))
    int_grid = np.stack(np.meshgrid(*[np.arange(2 ** Ls[d]) / 2.0 ** Ls[d] for d in dims], indexing="ij"),
                        -1).reshape(-1, len(dims))
    xs = np.zeros((rest_grid.shape[0] * int_grid.shape[0], s.dimension()))
    for k, d in enumerate(rest):
        xs[:, d - 1] = np.repeat(rest_grid[:, k], int_grid.shape[0])
    for k, d in enumerate(dims):
        xs[:, d - 1] = np.tile(int_grid[:, k], rest_grid.shape[0])
    y = oracle_eval(po, f, xs).reshape(rest_grid.shape[0], int_grid.shape[0])
    w = 1.0 if take_sum else 1.0 / int_grid.shape[0]
    return rest, rest_grid, y.sum(axis=1) * w


@pytest.mark.parametrize("case", ["tooth2", "tooth1_root", "take_sum", "no_merge", "chain2d"])
def test_partial_integrate(q, po, eng, rng, case):
    if case == "chain2d":
        s = q.continuous_siteinds(q.named_grid((10, 1)), map_dimension=2)
    else:
        s = _comb2(q, 5)
    f = q.random_itn(s, link_space=4, rng=rng)
    dims = [1] if case == "tooth1_root" else [2]
    take_sum = case == "take_sum"
    merge = case != "no_merge"
    if case == "chain2d":
        merge = False  # the remaining digits interleave with the integrated ones: no merged layout
    p = q.partial_integrate(f, dims, merge_vertices=merge, take_sum=take_sum)
    rest, grid, ref = _brute_partial(po, f, dims, take_sum)
    xs = np.zeros((grid.shape[0], p.dimension()))
    for k, d in enumerate(rest):
        xs[:, d - 1] = grid[:, k]
    y = q.evaluate(p, xs)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), case
    if merge:
        assert p.topo.nv == s.ndigits(rest[0])  # only the remaining digits' vertices
    # integrating the rest as well gives the full integral
    tot = q.integrate(p, take_sum=take_sum)
    full = po.integrate(f.topo.parent, f.topo.sdim, f.cores, 1.0 if take_sum else 0.5)
    assert tot == pytest.approx(full, rel=1e-11, abs=1e-13 * abs(full) + 1e-300)


def test_partial_integrate_unit_bond_edge(q, po, eng, rng):
    """A product of functions of x and of y joins the teeth by a bond of 1: the
    integrated tooth still contributes its integral through that edge."""
    s = _comb2(q, 5)
    a = q.random_itn(q.reduced_indsnetworkmap(s, 1), link_space=3, rng=rng)
    b = q.random_itn(q.reduced_indsnetworkmap(s, 2), link_space=2, rng=rng)
    ab = q.multiply(a, b)
    assert ab.bonds()[ab.topo.id[(2, 1)]] == 1
    grid = (np.arange(32) / 32.0)[:, None]
    mb = np.mean(oracle_eval(po, b, np.hstack([np.zeros((32, 1)), grid])))
    ma = np.mean(oracle_eval(po, a, grid))
    p2 = q.partial_integrate(ab, [2])
    assert np.allclose(q.evaluate(p2, grid), oracle_eval(po, a, grid) * mb, rtol=1e-12, atol=1e-13)
    p1 = q.partial_integrate(ab, [1])
    xs = np.hstack([np.zeros((32, 1)), grid])
    assert np.allclose(q.evaluate(p1, xs), oracle_eval(po, b, xs) * ma, rtol=1e-12, atol=1e-13)


def test_partial_integrate_disconnected_rest(q, po, eng, rng):
    s = q.continuous_siteinds(q.named_comb_tree((3, 3)), [[(i, j) for j in range(1, 4)] for i in (1, 2, 3)])
    f = q.random_itn(s, link_space=3, rng=rng)
    with pytest.raises(NotImplementedError):
        q.partial_integrate(f, [2])  # teeth 1 and 3 meet only through tooth 2
    p = q.partial_integrate(f, [2], merge_vertices=False)
    rest, grid, ref = _brute_partial(po, f, [2])
    xs = np.zeros((grid.shape[0], 3))
    xs[:, 0], xs[:, 2] = grid[:, 0], grid[:, 1]
    assert np.allclose(q.evaluate(p, xs), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_fredholm_iteration(q, eng):
    """f(x) = e^x + int_0^1 x y f(y) dy by fixed-point iteration on a comb of
    two teeth (examples/fredholm_solver.jl:43-65), every step on the device,
    against the same iteration on the dense grid (left Riemann sums)."""
    L = 7
    s = _comb2(q, L)
    s1, s2 = q.reduced_indsnetworkmap(s, 1), q.reduced_indsnetworkmap(s, 2)
    psi = q.const_itn(s2)
    g = q.poly_itn(s1, [0, 1], dim=1) * q.poly_itn(s2, [0, 1], dim=2)
    c1, c2 = q.exp_itn(s1, dim=1), q.exp_itn(s2, dim=2)
    N = 2 ** L
    x = np.arange(N) / N
    dense = np.ones(N)
    dim = 2
    for _ in range(12):
        psi = psi * g
        psi = q.partial_integrate(psi, [dim])
        dim = 1 if dim == 2 else 2
        psi = psi + (c1 if dim == 1 else c2)
        dense = x * np.mean(x * dense) + np.exp(x)
        assert psi.maxlinkdim() <= 3 * 2 + 1
    xs = np.zeros((N, 2))
    xs[:, dim - 1] = x
    y = q.evaluate(psi, xs)
    assert np.allclose(y, dense, rtol=1e-12, atol=0)
    # and the continuum solution 3x/2 + e^x up to the grid's discretisation
    assert np.max(np.abs(y - (1.5 * x + np.exp(x)))) < 10.0 / N


# --------------------------------------------------------------------------- operator construction
def test_operator_proj_reference_cases(q, po, eng, rng):
    """test/test_operators.jl:148-165 (comb (4, 3), 1D) and
    test/test_integration.jl:15-26 (integral of exp(x)^2 through operator_proj)."""
    import math
    s = q.continuous_siteinds(q.named_comb_tree((4, 3)))
    gx = q.sin_itn(s, k=0.5 * math.pi)
    fx = q.cos_itn(s, k=0.25 * math.pi)
    O = q.operator_proj(fx)
    # the operator's cores: f's on the diagonal of (s_in, s_out)
    t = fx.topo
    for v, (A, Oc) in enumerate(zip(fx.cores, O.cores)):
        assert Oc.shape == (A.shape[0], A.shape[1], A.shape[1]) + A.shape[2:]
        for si in range(A.shape[1]):
            for so in range(A.shape[1]):
                assert np.array_equal(Oc[:, si, so], A[:, si] if si == so else np.zeros_like(A[:, si]))
    fg = q.operate(O, gx, cutoff=1e-14)
    sq = q.operate(O, fx, cutoff=1e-14)
    for x in [0.025, 0.1, 0.25, 0.625, 0.875]:
        assert q.evaluate(fg, x) == pytest.approx(math.sin(0.5 * math.pi * x) * math.cos(0.25 * math.pi * x), abs=1e-3)
        assert q.evaluate(sq, x) == pytest.approx(math.cos(0.25 * math.pi * x) ** 2, abs=1e-3)
    # against the oracle: the exact product's values at random points
    xs = rng.uniform(0, 1, (300, 1))
    ref = oracle_eval(po, fx, xs) * oracle_eval(po, gx, xs)
    fg_exact = q.operate(O, gx, truncate=False)
    assert np.allclose(q.evaluate(fg_exact, xs), ref, rtol=1e-12, atol=1e-13)
    # integral of exp(x)^2, L = 20 on a comb (2, 10)
    s20 = q.continuous_siteinds(q.named_comb_tree((2, 10)))
    e = q.exp_itn(s20)
    Oe = q.operator_proj(e)
    assert q.integrate(Oe, e) == pytest.approx(0.5 * (math.e ** 2 - 1), abs=1e-4)


def test_map_to_zeros_reference_cases(q, eng):
    """test/test_operators.jl:355-411: boundary operators in 1D and 2D on combs."""
    s = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    L = 6
    delta = 2.0 ** -L
    last = 1 - delta
    xs = [0.0, delta, 0.25, 0.5, 0.625, 0.875, last]
    f = q.poly_itn(s, [1.0, 0.5, 0.25])
    Zo = q.map_to_zero_operator(s, [0, last])
    for p1 in (0.0, last):
        p = q.delta_p(s, p1)
        assert q.inner(p, p, op=Zo) == pytest.approx(0.0, abs=1e-14)
    phi = q.map_to_zeros(f, [0, last], cutoff=1e-16, maxdim=10)
    for x in xs:
        val = q.evaluate(phi, x)
        if x in (0.0, last):
            assert val == pytest.approx(0.0, abs=1e-8)
        else:
            assert val == pytest.approx(q.evaluate(f, x), rel=1e-10)
    # 2D comb (3, 3), interleaved
    s2 = q.continuous_siteinds(q.named_comb_tree((3, 3)), map_dimension=2)
    Ly = s2.ndigits(2)
    ly = 1 - 2.0 ** -Ly
    lx = 1 - 2.0 ** -s2.ndigits(1)
    ys = [0.0, 2.0 ** -Ly, 0.25, 0.5, 0.625, 0.875, ly]
    fxy = q.add(q.poly_itn(s2, [1.0, 0.5, 0.25], dim=1), q.cos_itn(s2, dim=2))
    Zo2 = q.map_to_zero_operator(s2, [0, lx, 0, ly], [1, 1, 2, 2])
    for p1 in (0.0, lx):
        for p2 in (0.0, ly):
            p = q.delta_p(s2, [p1, p2])
            assert q.inner(p, p, op=Zo2) == pytest.approx(0.0, abs=1e-14)
    phi2 = q.operate(Zo2, fxy, cutoff=0.0, maxdim=10)
    for x in (0.0, lx):
        vals = [q.evaluate(phi2, [x, y]) for y in ys]
        assert np.allclose(vals, 0.0, atol=1e-8)
    # interior points keep the function
    assert q.evaluate(phi2, [0.5, 0.25]) == pytest.approx(q.evaluate(fxy, [0.5, 0.25]), rel=1e-10)

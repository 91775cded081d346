"""CPU checks of the function-algebra support: the oracle's Hadamard product
and direct sum pinned by pointwise values (product and sum of the operands'
values at random points), the polynomial builder pinned by numpy's
polynomial values on the grid, and the layouts of reduced, merged and
integrated functions (src/indsnetworkmap.jl:29-46,
src/IndexMaps/abstractindexmap.jl:72-73, src/elementary_operators.jl:261)."""
import numpy as np
import pytest

from helpers import grid_of, oracle_eval


def _layouts(q):
    return {
        "chain": q.continuous_siteinds(q.named_grid((9, 1))),
        "chain2d": q.continuous_siteinds(q.named_grid((10, 1)), map_dimension=2),
        "comb": q.continuous_siteinds(q.named_comb_tree((3, 4)), [[(i, j) for j in range(1, 5)] for i in (1, 2, 3)]),
    }


@pytest.mark.parametrize("case", ["chain", "chain2d", "comb"])
def test_oracle_hadamard_and_sum_pointwise(q, po, rng, case):
    s = _layouts(q)[case]
    f = q.random_itn(s, link_space=3, rng=rng)
    g = q.random_itn(s, link_space=2, rng=rng)
    t = f.topo
    xs = rng.uniform(0.0, 1.0, (200, s.dimension()))
    yf, yg = oracle_eval(po, f, xs), oracle_eval(po, g, xs)
    H = po.hadamard(t.parent, t.sdim, g.cores, f.cores)
    S = po.direct_sum(t.parent, t.sdim, f.cores, g.cores)
    yh, rc = po.evaluate(t.parent, t.sdim, H, t.vcoord, t.vdigit, t.Lc, xs)
    assert rc == 0
    ys, rc = po.evaluate(t.parent, t.sdim, S, t.vcoord, t.vdigit, t.Lc, xs)
    assert rc == 0
    assert np.allclose(yh, yf * yg, rtol=1e-12, atol=1e-12 * np.abs(yf * yg).max())
    assert np.allclose(ys, yf + yg, rtol=1e-12, atol=1e-12 * np.abs(yf + yg).max())
    # bonds multiply / add (root 1)
    assert [c.shape[0] for c in H] == list(f.bonds() * g.bonds())
    sb = f.bonds() + g.bonds()
    sb[0] = 1
    assert [c.shape[0] for c in S] == list(sb)


@pytest.mark.parametrize("case", ["chain", "chain2d", "comb"])
def test_poly_itn_values(q, po, rng, case):
    s = _layouts(q)[case]
    coeffs = [0.3, -1.2, 0.7, 2.5]
    for dim in s.dimensions():
        f = q.poly_itn(s, coeffs, dim=dim, k=1.5, c=-0.8)
        assert f.maxlinkdim() == len(coeffs)
        xs = rng.uniform(0.0, 1.0, (100, s.dimension()))
        xg = grid_of(xs[:, dim - 1], s.ndigits(dim))
        ref = -0.8 * np.polynomial.polynomial.polyval(1.5 * xg, coeffs)
        assert np.allclose(oracle_eval(po, f, xs), ref, rtol=1e-12, atol=1e-12)
    # one coefficient: the constant
    f = q.poly_itn(s, [2.0])
    assert np.allclose(oracle_eval(po, f, rng.uniform(0, 1, (5, s.dimension()))), 2.0)


def test_reduced_and_union_layouts(q):
    s = q.continuous_siteinds(q.named_comb_tree((2, 5)), [[(i, j) for j in range(1, 6)] for i in (1, 2)])
    s1, s2 = q.reduced_indsnetworkmap(s, 1), q.reduced_indsnetworkmap(s, 2)
    assert s1.graph.vertices == [(1, j) for j in range(1, 6)]
    assert s1.dimension() == 1 and s1.ndigits(1) == 5
    # dimension 2 keeps its number; dimension 1 has no digit there
    assert s2.dimension() == 2 and s2.ndigits(1) == 0 and s2.ndigits(2) == 5
    t2 = q.Topology(s2)
    assert list(t2.Lc) == [0, 5] and t2.ncoord == 2
    # parts of one layout join through its own edges
    u = s1.union(s2)
    assert u.same_layout(s)
    assert q.Topology(u).same(q.Topology(s))
    # integrated digits removed, vertices kept
    w = s.without_dimensions([2])
    assert w.graph.vertices == s.graph.vertices and w.dimension() == 1
    # a layout whose remaining vertices are not connected
    s3 = q.continuous_siteinds(q.named_comb_tree((3, 3)), [[(i, j) for j in range(1, 4)] for i in (1, 2, 3)])
    with pytest.raises(NotImplementedError):
        s3.reduced([1, 3])
    # different digit placement cannot be merged
    a = q.continuous_siteinds(q.named_grid((4, 1)))
    b = q.continuous_siteinds(q.named_grid((4, 1)), [[(4, 1), (3, 1), (2, 1), (1, 1)]])
    with pytest.raises(ValueError):
        a.union(b)


def test_delta_p_and_kernel_values(q, po, rng):
    """delta_p is the indicator of its grid points; delta_kernel with the
    identity and coeff -1 vanishes exactly on the given grid lines and is 1
    elsewhere (the crossings of lines in different dimensions added back),
    as the reference's boundary tests expect (test/test_operators.jl:355-411)."""
    # 1D on a comb (2, 3): points 0 and the last grid point
    s = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    L = 6
    last = 1 - 2.0 ** -L
    grid = np.arange(2 ** L)[:, None] / 2.0 ** L
    k = q.delta_kernel(s, [[0.0], [last]], [[1], [1]])
    y = oracle_eval(po, k, grid)
    ref = np.ones(2 ** L)
    ref[[0, -1]] = 0.0
    assert np.allclose(y, ref, atol=1e-14)
    d = q.delta_p(s, 0.25)
    yd = oracle_eval(po, d, grid)
    assert np.allclose(yd, (np.arange(2 ** L) == 16).astype(float), atol=0)
    # 2D on a comb (3, 3), interleaved: lines x in {0, last}, y in {0, last}
    s2 = q.continuous_siteinds(q.named_comb_tree((3, 3)), map_dimension=2)
    Lx, Ly = s2.ndigits(1), s2.ndigits(2)
    lx, ly = 1 - 2.0 ** -Lx, 1 - 2.0 ** -Ly
    k2 = q.delta_kernel(s2, [[0.0], [lx], [0.0], [ly]], [[1], [1], [2], [2]])
    gx = np.repeat(np.arange(2 ** Lx) / 2.0 ** Lx, 2 ** Ly)
    gy = np.tile(np.arange(2 ** Ly) / 2.0 ** Ly, 2 ** Lx)
    y2 = oracle_eval(po, k2, np.stack([gx, gy], 1))
    on_x = np.isin(gx, [0.0, lx])
    on_y = np.isin(gy, [0.0, ly])
    assert np.allclose(y2, (~on_x & ~on_y).astype(float), atol=1e-13)
    # a repeated point is removed once, not twice
    k3 = q.delta_kernel(s, [[0.25], [0.25]], [[1], [1]])
    y3 = oracle_eval(po, k3, grid)
    assert y3[16] == pytest.approx(0.0, abs=1e-14) and np.allclose(np.delete(y3, 16), 1.0)
    # a 2D point (both coordinates) and the plane of the first coordinate
    pt = q.delta_p(s2, [[0.5, 0.25], [0.75, 0.5]], [[1, 2], [1, 2]])
    yp = oracle_eval(po, pt, np.array([[0.5, 0.25], [0.75, 0.5], [0.5, 0.5], [0.25, 0.25]]))
    assert np.allclose(yp, [1, 1, 0, 0])


def test_point_to_opsum_projects(q, po, rng):
    """The compiled operator of point_to_opsum keeps f at x's digits of the
    dimension and zeroes it elsewhere (src/elementary_operators.jl:200-210)."""
    s = q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
    f = q.random_itn(s, link_space=3, rng=rng)
    op = q.compile_terms(s, q.point_to_opsum(s, 0.375, 1))
    t = f.topo
    g = po.contract(t.parent, t.sdim, op.cores, f.cores)
    xs = rng.uniform(0, 1, (200, 2))
    xs[:50, 0] = 0.375
    y, _ = po.evaluate(t.parent, t.sdim, g, t.vcoord, t.vdigit, t.Lc, xs)
    yf = oracle_eval(po, f, xs)
    hit = grid_of(xs[:, 0], s.ndigits(1)) == 0.375
    assert np.allclose(y[hit], yf[hit], rtol=1e-12) and np.allclose(y[~hit], 0.0, atol=1e-14)

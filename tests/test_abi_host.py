"""CPU-side checks: the C-ABI library loads and exports every declared symbol,
and the host-side logic (digit layout, topology, builders, operators) behaves
as the reference's host code.  No engine compute runs here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from helpers import oracle_eval

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qtt.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qtt_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding(q):
    from itensornumericalanalysis_jl_amd import _capi
    assert sorted(_capi.EXPORTS) == _declared()


def test_library_exports_every_symbol(q):
    from itensornumericalanalysis_jl_amd import _capi
    if not os.path.exists(_capi.LIBPATH):
        pytest.fail("libqtt.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_capi.LIBPATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.qtt_version() == 200


def test_engine_refuses_without_device(q):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    s = q.continuous_siteinds(q.named_grid((4, 1)))
    f = q.exp_itn(s)
    from itensornumericalanalysis_jl_amd._capi import EngineError
    with pytest.raises(EngineError):
        q.evaluate(f, 0.5)


# ---------------------------------------------------------------- digit layout
def test_default_dimension_vertices(q):
    # src/digit_inds.jl:17-21: interleaved
    g = q.named_grid((6, 1))
    dv = q.default_dimension_vertices(g, map_dimension=2)
    assert dv == [[(1, 1), (3, 1), (5, 1)], [(2, 1), (4, 1), (6, 1)]]


def test_named_graph_orders(q):
    g = q.named_comb_tree((3, 2))
    assert g.vertices == [(1, 1), (2, 1), (3, 1), (1, 2), (2, 2), (3, 2)]
    assert g.is_tree()
    assert not q.named_grid((2, 2)).is_tree()


def test_loopy_graph_rejected(q):
    s = q.continuous_siteinds(q.named_grid((3, 3)))
    with pytest.raises(NotImplementedError):
        q.exp_itn(s)


def test_host_digit_map(q, po):
    s = q.continuous_siteinds(q.named_grid((12, 1)), map_dimension=2)
    vals = s.calculate_ind_values([0.625, 0.3], [1, 2])
    assert s.calculate_p(vals) == [0.625, math.floor(0.3 * 64) / 64]
    for x in (0.0, 0.1, 0.999, 1.0, 3.0):
        assert q.IndsNetworkMap.digits_of(x, 9) == [po.digit(x, 9, j) for j in range(1, 10)]
    with pytest.raises(q.DigitMapError):
        s.calculate_ind_values([-0.1, 0.2])
    assert s.grid_points(4, 1) == [0.0, 0.25, 0.5, 0.75]


def test_topology_bfs(q):
    s = q.continuous_siteinds(q.named_comb_tree((3, 4)), [[(i, j) for j in range(1, 5)] for i in (1, 2, 3)])
    t = q.Topology(s)
    assert t.parent[0] == -1
    assert all(0 <= t.parent[v] < v for v in range(1, t.nv))
    assert list(t.Lc) == [4, 4, 4]


# ---------------------------------------------------------------- builders / operators
def test_direct_sum_and_product_values(q, po, rng):
    s = q.continuous_siteinds(q.named_comb_tree((2, 4)), map_dimension=2)
    f = q.sin_itn(s, k=3.0, dim=1)
    g = q.exp_itn(s, k=-0.5, a=0.2, dim=2)
    xs = rng.uniform(0, 1, (50, 2))
    yf, yg = oracle_eval(po, f, xs), oracle_eval(po, g, xs)
    t = f.topo
    S = po.direct_sum(t.parent, t.sdim, f.cores, g.cores)
    H = po.hadamard(t.parent, t.sdim, f.cores, g.cores)
    ys, _ = po.evaluate(t.parent, t.sdim, S, t.vcoord, t.vdigit, t.Lc, xs)
    yh, _ = po.evaluate(t.parent, t.sdim, H, t.vcoord, t.vdigit, t.Lc, xs)
    assert np.allclose(ys, yf + yg, rtol=1e-13, atol=1e-14)
    assert np.allclose(yh, yf * yg, rtol=1e-13, atol=1e-14)
    assert np.allclose(oracle_eval(po, 2.5 * f, xs), 2.5 * yf, rtol=1e-14)
    assert max(c.shape[0] for c in S) == 3 and max(c.shape[0] for c in H) == 2


def test_operator_bonds_minimal(q):
    s = q.continuous_siteinds(q.named_grid((20, 1)))
    assert q.second_derivative_operator(s).maxlinkdim() == 3
    assert q.first_derivative_operator(s).maxlinkdim() == 3
    s2 = q.continuous_siteinds(q.named_grid((40, 1)), map_dimension=2)
    assert q.laplacian_operator(s2).maxlinkdim() <= 6


def test_workloads_shapes(q):
    from itensornumericalanalysis_jl_amd import workloads
    c1 = workloads.sine_1d(npts=100)
    assert c1.f.topo.nv == 12 and c1.f.maxlinkdim() == 5 and c1.xs.shape == (100, 1)
    c2 = workloads.laplacian_2d(npts=10)
    assert c2.f.topo.nv == 40 and c2.f.maxlinkdim() == 16 and c2.xs.shape == (10, 2)

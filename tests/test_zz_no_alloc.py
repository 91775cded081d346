"""qtt_eval_noalloc (evaluation staged in the library's fixed device array)
against qtt_eval through the C-ABI: two calls in a row, the second on a smaller
network and fewer points, plus the limits it must refuse."""
import ctypes
import math

import numpy as np
import pytest

from helpers import oracle_eval, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng(q):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return q.engine_context()


def _noalloc(q, eng, f, xd):
    import torch
    from itensornumericalanalysis_jl_amd import _capi
    dn = q.api.device_network(f)
    out = torch.full((xd.shape[0],), float("nan"), dtype=torch.float64, device="cuda")
    rc = _capi.lib().qtt_eval_noalloc(eng.handle, dn.handle, ctypes.c_void_p(xd.data_ptr()), xd.shape[0],
                                      ctypes.c_void_p(out.data_ptr()))
    eng.check(rc)
    eng.sync()
    return out.cpu().numpy()


def test_noalloc_matches_eval_two_calls(q, po, eng, rng):
    import torch
    cases = []
    # 1: a 40-site chain of bond 16 at 300001 points (two points per lane is
    #    not taken below 2^19 points: one-point register kernel)
    s1 = q.continuous_siteinds(q.named_grid((40, 1)), map_dimension=2)
    f1 = q.random_itn(s1, link_space=16, rng=rng, std=0.25)
    cases.append((f1, rng.uniform(0, 1, (300001, 2))))
    # 2: smaller network, fewer points (the array keeps the first call's tail)
    s2 = q.continuous_siteinds(q.named_grid((12, 1)))
    f2 = q.random_itn(s2, link_space=5, rng=rng, std=0.5)
    cases.append((f2, rng.uniform(0, 1, (1000, 1))))
    for f, xs in cases:
        xd = torch.tensor(xs, device="cuda")
        ref = q.evaluate(f, xd).cpu().numpy()
        got = _noalloc(q, eng, f, xd)
        # the same chain kernels; qtt_eval may also take environment tables
        # for many points, which changes only the rounding order
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        assert rel_l2(got, oracle_eval(po, f, xs)) < 1e-12


def test_noalloc_lds_kernel_and_many_points(q, po, eng, rng):
    import torch
    # bond 40: the LDS-vector chain kernel; then bond 8 at 2^19 + 7 points
    # (two points per lane), then the first again
    s = q.continuous_siteinds(q.named_grid((14, 1)), map_dimension=2)
    fa = q.random_itn(s, link_space=40, rng=rng, std=1.0 / math.sqrt(40))
    fb = q.random_itn(s, link_space=8, rng=rng, std=0.3)
    xa = rng.uniform(0, 1, (5000, 2))
    xb = rng.uniform(0, 1, ((1 << 19) + 7, 2))
    for f, xs in ((fa, xa), (fb, xb), (fa, xa[:100])):
        xd = torch.tensor(xs, device="cuda")
        ref = q.evaluate(f, xd).cpu().numpy()
        got = _noalloc(q, eng, f, xd)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        assert rel_l2(got[:4000], oracle_eval(po, f, xs[:4000])) < 1e-12


def test_noalloc_refuses_beyond_limits(q, eng, rng):
    import torch
    from itensornumericalanalysis_jl_amd import _capi
    L = _capi.lib()
    xd = torch.zeros((4, 1), dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.float64, device="cuda")

    def call(f):
        dn = q.api.device_network(f)
        return L.qtt_eval_noalloc(eng.handle, dn.handle, ctypes.c_void_p(xd.data_ptr()), 4,
                                  ctypes.c_void_p(out.data_ptr()))
    # a tree
    st = q.continuous_siteinds(q.named_comb_tree((2, 3)))
    assert call(q.random_itn(st, link_space=2, rng=rng)) == _capi.QTT_EINVAL
    # packed cores beyond QTT_NOALLOC_PACK_DOUBLES (2^20): 80 sites of bond 120
    # take 80 * 2 * 120 * 120 = 2.3e6 doubles
    sb = q.continuous_siteinds(q.named_grid((80, 1)), map_dimension=2)
    big = q.random_itn(sb, link_space=120, rng=rng, std=0.05)
    assert max(big.bonds()) == 120
    assert call(big) == _capi.QTT_EINVAL
    assert b"QTT_NOALLOC_PACK_DOUBLES" in L.qtt_last_error(eng.handle)
    # the context still evaluates afterwards
    sm = q.continuous_siteinds(q.named_grid((6, 1)))
    assert q.evaluate(q.exp_itn(sm), 0.5) == pytest.approx(math.exp(0.5), rel=1e-13)

"""Pl_bins_noalloc (include/fftgal_ext.h): the allocation-free form of
Pl_bins, whose bin sums live in the library's fixed device scratch.  Called
through the C ABI and checked against Pl_bins on the same grid, on two calls
in a row (the second on a smaller grid, with fewer bins, so leftovers of the
first call in the scratch would show), and for its size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

dptr = C.POINTER(C.c_double)
lptr = C.POINTER(C.c_long)


def _dp(a):
    return a.ctypes.data_as(dptr)


def _noalloc(lib, g, los, dK):
    nk = lib.fftgal_NK(g.Ng, g.L, float(dK))
    K = np.full(nk, -1.0)
    N = np.full(nk, -1, dtype=np.int64)
    P = np.full((nk, 4), -1.0)
    l = np.array(los, dtype=np.float64)
    rc = lib.Pl_bins_noalloc(g.ptr, _dp(l), float(dK), _dp(K), N.ctypes.data_as(lptr), _dp(P))
    return rc, K, N, P


def _power_grid(fa, Ng, L, Np, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.random(Np) * L for _ in range(3))
    g = fa.Grid(Ng, L)
    c = fa.Catalog(x, y, z, L ** 3)
    g.p2k(c)
    g.P(c)
    return g, c


def _same(got, ref):
    """counts exact; K and P within FP64 rounding of the atomic bin sums"""
    rc, K, N, P = got
    Kr, Nr, Pr = ref
    assert rc == len(Kr)
    assert np.array_equal(N, Nr)
    ok = Nr > 0
    assert np.all(np.isnan(K[~ok])) and np.all(np.isnan(P[~ok]))
    assert np.allclose(K[ok], Kr[ok], rtol=1e-13, atol=0)
    scale = np.abs(Pr[ok]).max(axis=0)
    assert np.all(np.abs(P[ok] - Pr[ok]) <= 1e-12 * scale)


def test_noalloc_matches_Pl_bins_twice_then_smaller(torch_cuda, lib):
    import fftgal_amd as fa
    L = 400.0
    big, cb = _power_grid(fa, 64, L, 40000, 1)
    small, cs = _power_grid(fa, 32, L, 9000, 2)
    for g, dK, los in ((big, 2 * np.pi / L, (0, 0, 1)), (big, 2 * np.pi / L, (0, 0, 1)),
                       (small, 2 * np.pi / L, (1, 1, 0)), (small, 4 * np.pi / L, (0, 0, 1))):
        print(f"Pl_bins_noalloc Ng={g.Ng} dK={dK:.4g} los={los}", flush=True)
        got = _noalloc(lib, g, los, dK)
        assert got[0] > 0, fa.load().fftgal_last_error()
        _same(got, g.Pl_bins(los, dK))
    for g in (big, small):
        g.free()


def test_noalloc_rejects_more_bins_than_its_scratch(torch_cuda, lib):
    import fftgal_amd as fa
    L = 400.0
    g, c = _power_grid(fa, 32, L, 5000, 3)
    kF = 2 * np.pi / L
    dK = kF / 800            # NK = floor(sqrt(3) * 16 * 800) + 1 = 22171 > 21840
    assert lib.fftgal_NK(32, L, dK) == 22171
    rc, K, N, P = _noalloc(lib, g, (0, 0, 1), dK)
    assert rc == -1
    msg = lib.fftgal_last_error()
    msg = msg.decode() if isinstance(msg, bytes) else str(msg)
    assert "exceed" in msg and "21840" in msg
    assert np.all(N == -1) and np.all(K == -1.0)      # outputs untouched
    dK = kF / 780            # NK = 21616: within the limit
    assert lib.fftgal_NK(32, L, dK) == 21616
    print("Pl_bins_noalloc at NK=21616", flush=True)
    _same(_noalloc(lib, g, (0, 0, 1), dK), g.Pl_bins((0, 0, 1), dK))
    g.free()

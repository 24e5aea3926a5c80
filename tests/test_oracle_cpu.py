"""Pins of the FP64 oracle (CPU).  Each check is independent of the
oracle's own formulas: the reference's known-answer test
(src/tests/octet.c), numpy.fft golden vectors, brute-force DFT, the PCS
kernel written as the cubic B-spline M4, full-grid enumeration of the
k-space bins with numpy's Legendre polynomials, and closed forms."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")
SIZES = (1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 16)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("Ng", SIZES)
def test_fft_against_numpy_golden(orc, gold, Ng):
    X = orc.fft_r2c_3d(gold[f"x_{Ng}"])
    ref = gold[f"X_{Ng}"]
    assert np.abs(X - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1)
    y = orc.fft_c2r_3d(gold[f"Y_{Ng}"], Ng)
    ref = gold[f"y_{Ng}"]
    assert np.abs(y - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1)


@pytest.mark.parametrize("Ng", [2, 3, 4])
def test_fft_brute_force_dft(orc, Ng):
    rng = np.random.default_rng(Ng)
    x = rng.standard_normal((Ng, Ng, Ng))
    n = np.arange(Ng)
    X = np.zeros((Ng, Ng, Ng // 2 + 1), dtype=complex)
    for i in range(Ng):
        for j in range(Ng):
            for k in range(Ng // 2 + 1):
                ph = np.exp(-2j * np.pi * (i * n[:, None, None] + j * n[None, :, None]
                                           + k * n[None, None, :]) / Ng)
                X[i, j, k] = (x * ph).sum()
    assert np.abs(orc.fft_r2c_3d(x) - X).max() <= 1e-13 * np.abs(X).max()


def test_fft_parseval_and_roundtrip(orc):
    Ng = 24
    rng = np.random.default_rng(9)
    x = rng.standard_normal((Ng, Ng, Ng))
    X = orc.fft_r2c_3d(x)
    w = np.full(X.shape[2], 2.0)
    w[0] = 1.0
    w[-1] = 1.0          # Ng even: k = Ng/2 counted once
    e_k = (np.abs(X) ** 2 * w).sum() / Ng ** 3
    assert abs(e_k - (x ** 2).sum()) <= 1e-12 * e_k
    back = orc.fft_c2r_3d(X, Ng) / Ng ** 3
    assert np.abs(back - x).max() <= 1e-13


def test_octet_known_answer(orc):
    """src/tests/octet.c:8-62: Ng = 2, L = 2, one particle at (-8,-8,-8),
    offset (7,7,7) -> -1; eps = 1e-15."""
    eps = 1e-15
    g = orc.Grid(2, 2.0)
    c = orc.Catalog([-8.0], [-8.0], [-8.0], 8.0)
    off = g.p2g(c, (7, 7, 7))
    assert list(off) == [-1.0, -1.0, -1.0]
    ijk = np.indices((2, 2, 2))
    exp = 8.0 / 27 * (ijk[0] + 1) * (ijk[1] + 1) * (ijk[2] + 1)
    assert np.abs(g.f[:, :, :2] - exp).max() < eps
    g.x2k(1)
    exp = 8.0 / (1 + 2 * ijk[0][:, :, 0]) / (1 + 2 * ijk[1][:, :, 0])
    assert np.abs(g.f[:, :, 0] - exp).max() < eps
    assert np.abs(g.f[:, :, 1]).max() < eps
    g.k2x(0)
    exp = 64.0 / 27 / (ijk[0] + 1) / (ijk[1] + 1) / (ijk[2] + 1)
    assert np.abs(g.f[:, :, :2] - exp).max() < eps


def m4(s):
    """cubic B-spline (PCS kernel) by its piecewise definition"""
    s = np.abs(s)
    return np.where(s < 1, (4 - 6 * s ** 2 + 3 * s ** 3) / 6,
                    np.where(s < 2, (2 - s) ** 3 / 6, 0.0))


def test_paint_single_particle_is_the_b_spline(orc):
    Ng, L = 8, 16.0
    rng = np.random.default_rng(1)
    for _ in range(20):
        pos = rng.random(3) * 3 * L - L
        off = rng.random(3) * 4 - 2
        g = orc.Grid(Ng, L)
        c = orc.Catalog([pos[0]], [pos[1]], [pos[2]], L ** 3)
        o = g.p2g(c, off)
        u = pos * Ng / L - o
        exp = np.zeros((Ng, Ng, Ng))
        n = np.arange(-3 * Ng, 4 * Ng)   # every periodic image of every node
        w = [m4(u[a] - n) for a in range(3)]
        for a in range(3):
            w[a] = np.bincount(n % Ng, weights=w[a], minlength=Ng)
        exp = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
        exp *= Ng ** 3      # 1 / (nbar H^3) with Np = 1, V = L^3
        assert np.abs(g.f[:, :, :Ng] - exp).max() <= 1e-12


def test_paint_mass_fold_and_translation(orc):
    Ng, L, fold = 16, 100.0, 2
    rng = np.random.default_rng(4)
    x, y, z = rng.random((3, 2000)) * L
    V = L ** 3
    # folding: grid box L/fold, catalog volume still L^3 -> mean fold^3
    g = orc.Grid(Ng, L / fold)
    c = orc.Catalog(x, y, z, V)
    g.p2g(c, (0, 0, 0))
    assert abs(g.f[:, :, :Ng].mean() - fold ** 3) <= 1e-12 * fold ** 3
    # one-cell translation along x is a cyclic shift of the grid
    H = L / fold / Ng
    g2 = orc.Grid(Ng, L / fold)
    c2 = orc.Catalog(x + H, y, z, V)
    g2.p2g(c2, (0, 0, 0))
    assert np.abs(np.roll(g.f[:, :, :Ng], 1, axis=0) - g2.f[:, :, :Ng]).max() <= 1e-10


def test_offset_phase_closed_form(orc):
    Ng, L = 6, 1.0
    g = orc.Grid(Ng, L)
    g.p2g(orc.Catalog([0.1], [0.2], [0.3], 1.0), (0.25, -0.5, 1.25))
    o = np.array(g.offset)
    g.f[:] = 0
    g.f[:, :, 0::2] = 1.0           # every mode = 1 + 0i
    g.offset_phase(1)
    kv = np.where(np.arange(Ng) * 2 < Ng, np.arange(Ng), np.arange(Ng) - Ng)
    kz = kv[: Ng // 2 + 1]
    exp = np.exp(-2j * np.pi * (o[0] * kv[:, None, None] + o[1] * kv[None, :, None]
                                + o[2] * kz[None, None, :]) / Ng)
    assert np.abs(g.fk() - exp).max() <= 1e-14
    g.offset_phase(-1)
    assert np.abs(g.fk() - 1).max() <= 1e-14


def test_deconv_closed_form(orc):
    Ng = 10
    g = orc.Grid(Ng, 5.0)
    g.f[:] = 0
    g.f[:, :, 0::2] = 1.0
    g.deconv()
    a = np.pi * np.abs(np.where(np.arange(Ng) * 2 <= Ng, np.arange(Ng), Ng - np.arange(Ng))) / Ng
    w = np.ones(Ng)
    w[1:] = (a[1:] / np.sin(a[1:])) ** 4
    exp = w[:, None, None] * w[None, :, None] * w[None, None, : Ng // 2 + 1]
    assert np.abs(g.fk().real - exp).max() <= 1e-13 * exp.max()


def full_grid_bins(Pk_half, Ng, L, los, dK, Nmu):
    """Bins by enumerating modes one by one (weight 1 each, DC excluded):
    every stored mode f = (Kval_i, Kval_j, Kval_k) of the half spectrum and,
    for 0 < k < Ng/2, its partner -f carrying the same P (Hermitian
    symmetry).  Kval gives the Nyquist index -Ng/2 (src/power.c:46-51); with
    a line of sight along a Nyquist axis this pairing, not the index grid,
    defines mu of the partner.  Legendre polynomials from numpy."""
    from numpy.polynomial import legendre
    kv = np.where(2 * np.arange(Ng) < Ng, np.arange(Ng), np.arange(Ng) - Ng).astype(float)
    kh = kv[: Ng // 2 + 1]
    FX, FY, FZ = np.meshgrid(kv, kv, kh, indexing="ij")
    inner = (np.arange(Ng // 2 + 1) > 0) & (2 * np.arange(Ng // 2 + 1) != Ng)
    sel = np.broadcast_to(inner[None, None, :], FX.shape)
    KX = np.concatenate([FX.ravel(), -FX[sel]])
    KY = np.concatenate([FY.ravel(), -FY[sel]])
    KZ = np.concatenate([FZ.ravel(), -FZ[sel]])
    P = np.concatenate([Pk_half.ravel(), Pk_half[sel]])
    kamp = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    # mode counts independently from the full index grid (fftfreq)
    ff = np.fft.fftfreq(Ng, 1.0 / Ng)
    kfull = np.sqrt(ff[:, None, None] ** 2 + ff[None, :, None] ** 2 + ff[None, None, :] ** 2)
    assert np.array_equal(np.sort(kfull.ravel()), np.sort(kamp))
    l = np.array(los, float)
    nl = np.linalg.norm(l)
    if nl > 1e-7:
        l = l / nl
    m = kamp > 0
    mu = np.zeros_like(kamp)
    mu[m] = (KX * l[0] + KY * l[1] + KZ * l[2])[m] / kamp[m]
    KF = 2 * np.pi / L
    ib = np.floor(kamp * KF / dK).astype(int)
    NK = int(np.floor(np.sqrt(3) * (Ng // 2) * KF / dK)) + 1
    N = np.bincount(ib[m], minlength=NK)
    K = np.bincount(ib[m], weights=kamp[m], minlength=NK)
    Pl = np.zeros((NK, 4))
    for q in range(4):
        c = np.zeros(2 * q + 1)
        c[-1] = 1
        Lq = legendre.legval(mu[m], c)
        Pl[:, q] = np.bincount(ib[m], weights=P[m] * Lq, minlength=NK)
    with np.errstate(invalid="ignore", divide="ignore"):
        Kout = K * KF / N
        Pl = Pl * (4 * np.arange(4) + 1) / N[:, None]
    im = np.minimum(np.floor(np.abs(mu) * Nmu).astype(int), Nmu - 1)
    Nw = np.zeros((NK, Nmu), dtype=np.int64)
    Pw = np.zeros((NK, Nmu))
    np.add.at(Nw, (ib[m], im[m]), 1)
    np.add.at(Pw, (ib[m], im[m]), P[m])
    with np.errstate(invalid="ignore", divide="ignore"):
        Pw = np.where(Nw > 0, Pw / np.maximum(Nw, 1), 0.0)
    return Kout, N, Pl, Nw, Pw


@pytest.mark.parametrize("Ng,los,dKf", [(8, (0, 0, 1), 1.0), (10, (1, 1, 0), 1.0),
                                        (12, (1, 2, 3), 0.7), (16, (0, 0, 0), 2.0)])
def test_P_Pl_Pmu_against_full_grid_enumeration(orc, Ng, los, dKf):
    L = 50.0
    rng = np.random.default_rng(Ng)
    x, y, z = rng.random((3, 300)) * L
    c = orc.Catalog(x, y, z, L ** 3)
    g = orc.Grid(Ng, L)
    g.p2k(c)
    fk = g.fk()
    g.P(c)
    # P closed form (src/power.c:12-27)
    Pexp = np.abs(fk) ** 2 / L ** 3 - L ** 3 / 300
    Pexp[0, 0, 0] = 0
    assert np.abs(g.fk().real - Pexp).max() <= 1e-12 * np.abs(Pexp).max()
    assert np.abs(g.fk().imag).max() == 0
    dK = dKf * 2 * np.pi / L
    Nmu = 3
    K, N, Pl = g.Pl_bins(los, dK)
    Km, Nm, Pm = g.Pmu_bins(los, dK, Nmu)
    Kr, Nr, Plr, Nwr, Pwr = full_grid_bins(g.fk().real, Ng, L, los, dK, Nmu)
    assert np.array_equal(N, Nr) and N.sum() == Ng ** 3 - 1
    ok = N > 0
    assert np.allclose(K[ok], Kr[ok], rtol=1e-13)
    assert np.all(np.isnan(Pl[~ok]))          # empty bins print NaN (0/0)
    assert np.allclose(Pl[ok], Plr[ok], rtol=1e-9, atol=1e-9 * np.abs(Plr[ok]).max())
    assert np.array_equal(Nm, Nwr) and Nm.sum() == Ng ** 3 - 1
    assert np.allclose(Pm, Pwr, rtol=1e-9, atol=1e-9 * np.abs(Pwr).max())


def test_pl_of_unit_power_is_closed_form(orc):
    """P(k) = 1 everywhere, los = 0 (mu = 0): P0 = 1, P2 = 5 L2(0) = -2.5,
    P4 = 9 L4(0) = 3.375, P6 = 13 L6(0) = -4.0625 in every non-empty bin."""
    Ng, L = 12, 30.0
    g = orc.Grid(Ng, L)
    g.f[:] = 0
    g.f[:, :, 0::2] = 1.0
    K, N, Pl = g.Pl_bins((0, 0, 0), 2 * np.pi / L)
    ok = N > 0
    assert np.allclose(Pl[ok], [1.0, -2.5, 3.375, -4.0625], rtol=1e-14)


def test_gal_rsd_closed_form(orc):
    L = 100.0
    rng = np.random.default_rng(8)
    x, y, z = rng.random((3, 50)) * L
    v = rng.standard_normal((3, 50)) * 200
    c = orc.Catalog(x, y, z, L ** 3, *v)
    los = np.array([0.0, 3.0, 4.0])
    aH = 70.0
    r = orc.load().oracle_gal_rsd(c.ptr, los.ctypes.data_as(orc.oracle.dptr), aH)
    lh = np.array([0, 0.6, 0.8])
    assert np.allclose(los, lh, rtol=0, atol=1e-16)
    s = (lh @ v) / aH
    for a, (name, p0) in enumerate((("x", x), ("y", y), ("z", z))):
        got = np.ctypeslib.as_array(getattr(r.contents, name), shape=(50,))
        assert np.allclose(got, p0 + s * lh[a], rtol=0, atol=1e-12)
    orc.load().oracle_gal_free(r)


def _dl(orc, g, kind, Nsub, los, ell):
    import ctypes as C
    l = np.array(los, dtype=np.float64)
    out = np.zeros(Nsub ** 3)
    dp = C.POINTER(C.c_double)
    orc.load().oracle_dl_filter(g.ptr, kind, Nsub, l.ctypes.data_as(dp), ell, 1.0,
                                out.ctypes.data_as(dp))
    return out.reshape(Nsub, Nsub, Nsub)


def test_dl_subbox_monopole_is_the_subbox_mean(orc):
    """src/app/DLsb.c: the sub-box window W(k) = mean over s < Ng/Nsub of
    e^{2 pi i k s / Ng} per axis, so the monopole at a corner is the mean of
    the field (less its grid mean, the DC mode being zeroed) over the sub box
    starting at that corner."""
    Ng, Nsub, L = 24, 4, 100.0
    rng = np.random.default_rng(3)
    x = rng.standard_normal((Ng, Ng, Ng))
    g = orc.Grid(Ng, L)
    g.f[:, :, :Ng] = x
    g.x2k(0)
    got = _dl(orc, g, 0, Nsub, (0, 0, 1), 0)
    s = Ng // Nsub
    ref = (x - x.mean()).reshape(Nsub, s, Nsub, s, Nsub, s).mean(axis=(1, 3, 5))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    g.free()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("ell", [0, 2, 4])
def test_dl_plane_wave(orc, kind, ell):
    """One plane wave cos(2 pi n.x / Ng) (n with n_z > 0): the filtered field
    at every corner is L_ell(mu_n) Re[W(n) e^{2 pi i n.x / Ng}], with W the
    sub-box window from its defining sum (kind 0) or the spherical top hat of
    radius sqrt(3) L / (2 Nsub) from a quadrature of its defining integral
    (kind 1), and L_ell from numpy."""
    from numpy.polynomial import legendre
    from scipy.integrate import quad
    Ng, Nsub, L = 24, 3, 300.0
    n = np.array([2, -1, 3])
    ax = np.arange(Ng)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    phase = 2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / Ng
    g = orc.Grid(Ng, L)
    g.f[:, :, :Ng] = np.cos(phase)
    g.x2k(0)
    los = np.array([1.0, 2.0, 2.0]) / 3.0
    got = _dl(orc, g, kind, Nsub, los, ell)
    s = Ng // Nsub
    if kind == 0:
        W = np.prod([np.exp(2j * np.pi * k * np.arange(s) / Ng).mean() for k in n % Ng])
    else:
        kR = np.linalg.norm(n) * np.pi * np.sqrt(3) / Nsub     # |k| R with R = sqrt(3) L / (2 Nsub)
        W = 3 * quad(lambda r: np.sin(kR * r) / (kR * r) * r * r, 0, 1, epsabs=1e-15)[0]
    mu = n @ los / np.linalg.norm(n)
    Lell = legendre.legval(mu, [0] * ell + [1])
    corners = phase[::s, ::s, ::s]
    ref = Lell * np.real(W * np.exp(1j * corners))
    assert np.abs(got - ref).max() <= 1e-12, (np.abs(got - ref).max())
    g.free()


_SUB_SCRIPT = r"""
import sys, json, ctypes as C
sys.path.insert(0, {root!r})
import numpy as np
import oracle, fftgal_amd as fa
o = oracle.load()
lib = fa.load()
rng = np.random.default_rng(1)
L = 100.0
x, y, z = rng.random((3, 4000)) * L
co = oracle.Catalog(x, y, z, L ** 3)
cg = fa.Catalog(x, y, z, L ** 3)
dp = C.POINTER(C.c_double)
out = []
for kind, vol, alpha in {calls!r}:
    v = np.array(vol, dtype=np.float64)
    if kind == "box":
        a = o.oracle_gal_subbox(co.ptr, v.ctypes.data_as(dp), alpha)
        b = lib.gal_subbox(cg.ptr, v.ctypes.data_as(dp), alpha)
    else:
        a = o.oracle_gal_subsphere(co.ptr, v.ctypes.data_as(dp), alpha)
        b = lib.gal_subsphere(cg.ptr, v.ctypes.data_as(dp), alpha)
    rec = {{}}
    for tag, s in (("orc", a.contents), ("lib", b.contents)):
        r = s.rand.contents
        rec[tag] = {{"Np": s.Np, "V": s.V, "Nr": r.Np, "Vr": r.V,
                    "x": [s.x[i] for i in range(s.Np)], "y": [s.y[i] for i in range(s.Np)],
                    "z": [s.z[i] for i in range(s.Np)],
                    "rx": [r.x[i] for i in range(r.Np)], "ry": [r.y[i] for i in range(r.Np)],
                    "rz": [r.z[i] for i in range(r.Np)]}}
    out.append(rec)
print(json.dumps({{"x": list(x), "y": list(y), "z": list(z), "calls": out}}))
"""


def _run_sub(calls, seed=None):
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ)
    env.pop("GSL_RNG_SEED", None)
    if seed is not None:
        env["GSL_RNG_SEED"] = str(seed)
    r = subprocess.run([sys.executable, "-c", _SUB_SCRIPT.format(root=root, calls=calls)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.mark.parametrize("seed", [None, 12345])
def test_subbox_subsphere_randoms_and_selection(seed):
    """gal_subbox / gal_subsphere (src/gal.c:78-158), library and oracle, in a
    fresh process each (their generators live for the process): selections
    against a numpy mask of the periodic sub-volume; randoms against numpy's
    MT19937 with the legacy integer seeding (GSL's seeding: GSL_RNG_SEED,
    0 -> 4357), u = output / 2^32, drawn x, y, z per random in sequence and
    continuing across calls; library and oracle bit-identical."""
    L = 100.0
    calls = [("box", [0.0, 50.0, 50.0, 100.0, 80.0, 120.0, L], 0.05),
             ("box", [10.0, 60.0, 0.0, 50.0, 0.0, 50.0, L], 0.1),
             ("sphere", [5.0, 95.0, 50.0, 20.0, L], 0.05),
             ("sphere", [50.0, 50.0, 50.0, 30.0, L], 0.08)]
    res = _run_sub(calls, seed)
    x, y, z = (np.array(res[k]) for k in "xyz")
    gens = {"box": np.random.RandomState(seed or 4357)._bit_generator,
            "sphere": np.random.RandomState(seed or 4357)._bit_generator}

    def u(kind):
        return float(gens[kind].random_raw()) / 4294967296.0

    for (kind, vol, alpha), rec in zip(calls, res["calls"]):
        assert rec["orc"] == rec["lib"], kind        # bit-identical, same order
        r = rec["lib"]
        if kind == "box":
            x0, x1, y0, y1, z0, z1, _ = vol
            x1, y1, z1 = (a0 + np.remainder(a1 - a0, L) if a1 - a0 != L else a1
                          for a0, a1 in ((x0, x1), (y0, y1), (z0, z1)))
            inside = ((np.mod(x - x0, L) < x1 - x0) & (np.mod(y - y0, L) < y1 - y0) &
                      (np.mod(z - z0, L) < z1 - z0))
            V = (x1 - x0) * (y1 - y0) * (z1 - z0)
        else:
            x0, y0, z0, R, _ = vol
            d = [np.remainder(a - a0, L) for a, a0 in ((x, x0), (y, y0), (z, z0))]
            d = [np.where(t >= L / 2, t - L, t) for t in d]
            inside = d[0] ** 2 + d[1] ** 2 + d[2] ** 2 < R * R
            V = 4 * np.pi / 3 * R ** 3
        assert np.array_equal(np.array(r["x"]), x[inside])
        assert np.array_equal(np.array(r["z"]), z[inside])
        assert r["Nr"] == round(inside.sum() / alpha) and r["V"] == r["Vr"]
        assert abs(r["V"] - V) <= 1e-12 * V
        for i in range(r["Nr"]):
            if kind == "box":
                exp = (x0 + (x1 - x0) * u(kind), y0 + (y1 - y0) * u(kind),
                       z0 + (z1 - z0) * u(kind))
            else:
                while True:
                    exp = (x0 + R * (2 * u(kind) - 1), y0 + R * (2 * u(kind) - 1),
                           z0 + R * (2 * u(kind) - 1))
                    dd = [np.remainder(e - e0, L) for e, e0 in zip(exp, (x0, y0, z0))]
                    dd = [t - L if t > L / 2 else t for t in dd]
                    if dd[0] ** 2 + dd[1] ** 2 + dd[2] ** 2 < R * R:
                        break
            assert (r["rx"][i], r["ry"][i], r["rz"][i]) == exp, (kind, i)


@pytest.mark.parametrize("kind,Ng,Nsub", [(0, 16, 2), (0, 24, 4), (1, 16, 4), (1, 24, 2)])
def test_subvolume_deconv_window(orc, kind, Ng, Nsub):
    """src/app/Psb.c:73-85 / Pss.c:70-87 on a field whose correlation
    function is 1 everywhere: after the deconvolution each lag holds the
    product of the inverse window autocorrelations of the lags that reach it
    -- for the box, the overlap count of a digitised gs^3 box with itself
    shifted by the lag (brute force); for the sphere, the lens volume of two
    spheres of diameter D = sqrt(3) gs at the lag's distance over the sphere
    volume, (4R + r)(2R - r)^2 / (16 R^3), for |lag|^2 < 0.99 D^2."""
    L = 50.0
    g = orc.Grid(Ng, L)
    g.f[:, :, :Ng] = 1.0
    g.x2k(0)
    orc.load().oracle_subvolume_deconv(g.ptr, kind, Nsub)
    g.k2x(0)
    got = g.f[:, :, :Ng].copy()
    gs = Ng // Nsub
    exp = np.ones((Ng, Ng, Ng))
    if kind == 0:
        box = np.zeros((Ng, Ng, Ng), bool)
        box[:gs, :gs, :gs] = True
        for i in range(-gs + 1, gs):
            for j in range(-gs + 1, gs):
                for k in range(-gs + 1, gs):
                    ov = (box & np.roll(box, (i, j, k), axis=(0, 1, 2))).sum()
                    exp[i % Ng, j % Ng, k % Ng] *= gs ** 3 / ov
    else:
        D = np.sqrt(3.0) * gs
        R = D / 2
        b = int(np.ceil(D))
        for i in range(-b + 1, b):
            for j in range(-b + 1, b):
                for k in range(-b + 1, b):
                    r2 = float(i * i + j * j + k * k)
                    if r2 < 0.99 * D * D:
                        r = np.sqrt(r2)
                        frac = (4 * R + r) * (2 * R - r) ** 2 / (16 * R ** 3)
                        exp[i % Ng, j % Ng, k % Ng] /= frac
    assert np.abs(got - exp).max() <= 1e-11 * np.abs(exp).max()
    g.free()


@pytest.mark.parametrize("Ng,interlace", [(48, True), (30, False), (2, True)])
def test_threaded_oracle_is_bit_identical(orc, Ng, interlace):
    """The oracle's threaded sweeps (paint by i-slabs, each filled in the
    serial particle order; whole 1D FFTs; k-space planes) give the serial
    result bit for bit, randoms included."""
    rng = np.random.default_rng(Ng)
    L = 300.0
    x, y, z = rng.random((3, 5000)) * L * 1.3 - 0.1 * L
    rx, ry, rz = rng.random((3, 9000)) * L
    out = []
    for threads in (1, 7):
        orc.set_threads(threads)
        try:
            r = orc.Catalog(rx, ry, rz, L ** 3)
            c = orc.Catalog(x, y, z, L ** 3, rand=r)
            g = orc.Grid(Ng, L)
            if interlace:
                g.p2k(c)
            else:
                g.p2g(c, (0.25, 0.0, 0.5))
                g.x2k(1)
                g.deconv()
            g.P(c)
            K, N, P = g.Pl_bins((0, 1, 1), 2 * np.pi / L)
            f = g.f.copy()
            g.k2x(0)
            out.append((f, g.f.copy(), K, N, P))
            g.free()
            c.free()
        finally:
            orc.set_threads(1)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f")

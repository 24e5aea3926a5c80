"""GPU parity of the drop-in API (include/fft.h, power.h) against the FP64
oracle, through the C library.  Mirrors the reference's own test
(src/tests/octet.c) and extends it to every hot-path function."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _catalogs(orc, fa, x, y, z, V, **kw):
    return orc.Catalog(x, y, z, V, **kw), fa.Catalog(x, y, z, V, **kw)


def test_octet_known_answer(torch_cuda, lib):
    """src/tests/octet.c:8-62, same inputs and eps = 1e-15."""
    import fftgal_amd as fa
    eps = 1e-15
    g = fa.Grid(2, 2.0)
    c = fa.Catalog([-8.0], [-8.0], [-8.0], 8.0)
    off = g.p2g(c, (7, 7, 7))
    assert list(off) == [-1.0, -1.0, -1.0] and g.offset == [-1.0, -1.0, -1.0]
    for i in range(2):
        for j in range(2):
            for k in range(2):
                assert abs(g.f[i, j, k] - 8.0 / 27 * (i + 1) * (j + 1) * (k + 1)) < eps
    g.x2k(1)
    for i in range(2):
        for j in range(2):
            exp = 8.0 / (1 + 2 * i) / (1 + 2 * j)
            assert abs(g.f[i, j, 0] - exp) < eps and abs(g.f[i, j, 1]) < eps
    g.k2x(0)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                assert abs(g.f[i, j, k] - 64.0 / 27 / (i + 1) / (j + 1) / (k + 1)) < eps


@pytest.mark.parametrize("Ng", [2, 4, 6, 8, 10, 12, 16, 24, 30, 48, 64, 96,
                                # any side, as FFTW plans it (src/fft.c:28-29): odd
                                # sides (paired real rows, Stockham), prime factors
                                # above 13 (Bluestein), and the smallest grids
                                1, 3, 5, 9, 15, 17, 21, 34, 38, 45, 97])
def test_x2k_k2x_vs_oracle(torch_cuda, lib, orc, Ng):
    import fftgal_amd as fa
    rng = np.random.default_rng(Ng)
    x = rng.standard_normal((Ng, Ng, Ng))
    L = 3.0 * Ng
    g = fa.Grid(Ng, L)
    o = orc.Grid(Ng, L)
    g.f[:, :, :Ng] = x
    o.f[:, :, :Ng] = x
    g.x2k(0)
    o.x2k(0)
    scale = np.abs(o.fk()).max()
    assert np.abs(g.fk() - o.fk()).max() <= 1e-13 * scale
    g.k2x(0)
    o.k2x(0)
    assert np.abs(g.f[:, :, :Ng] - o.f[:, :, :Ng]).max() <= 1e-13 * np.abs(x).max()
    assert np.abs(g.f[:, :, :Ng] - x).max() <= 1e-12 * np.abs(x).max()


def test_p2g_vs_oracle_with_randoms(torch_cuda, lib, orc):
    import fftgal_amd as fa
    rng = np.random.default_rng(7)
    Ng, L = 16, 100.0
    x, y, z = (rng.random((3, 3000)) * 2.3 - 0.6) * L   # outside [0, L) too: wrap
    xr, yr, zr = rng.random((3, 6000)) * L
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    ro = orc.Catalog(xr, yr, zr, L ** 3)
    co = orc.Catalog(x, y, z, L ** 3, rand=ro)
    # the library's gal_t with randoms
    cg = fa.Catalog(x, y, z, L ** 3)
    rg = fa.Catalog(xr, yr, zr, L ** 3)
    cg.ptr.contents.rand = rg.ptr
    rg.ptr = None   # now owned by the parent (gal_free frees rand)
    off = (0.3, -2.7, 21.5)
    oo = go.p2g(co, off)
    og = gg.p2g(cg, off)
    assert np.array_equal(oo, og) and go.offset == gg.offset
    ref = go.f[:, :, :Ng]
    assert np.abs(gg.f[:, :, :Ng] - ref).max() <= 1e-13 * np.abs(ref).max()


def test_p2g_empty_catalog_is_nan_like_the_reference(torch_cuda, lib, orc):
    """Np = 0: nothing is painted and the normalisation 1/(nbar H^3) with
    nbar = 0 (src/fft.c:125-129) turns the zero grid into NaN; the library
    launches nothing for the empty catalog and gives the same grid."""
    import fftgal_amd as fa
    Ng, L = 16, 100.0
    e = np.zeros(0)
    co, cg = _catalogs(orc, fa, e, e, e, L ** 3)
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    go.p2g(co, (0, 0, 0))
    gg.p2g(cg, (0, 0, 0))
    assert np.isnan(go.f[:, :, :Ng]).all()
    assert np.isnan(gg.f[:, :, :Ng]).all()


def test_offset_phase_roundtrip_and_deconv(torch_cuda, lib, orc):
    import fftgal_amd as fa
    Ng, L = 12, 50.0
    rng = np.random.default_rng(3)
    x, y, z = rng.random((3, 500)) * L
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    go.p2g(co, (0.25, 0.5, -0.75))
    gg.p2g(cg, (0.25, 0.5, -0.75))
    go.x2k(1)
    gg.x2k(1)
    s = np.abs(go.fk()).max()
    assert np.abs(gg.fk() - go.fk()).max() <= 1e-13 * s
    before = gg.fk()
    gg.offset_phase(1)
    gg.offset_phase(-1)
    assert np.abs(gg.fk() - before).max() <= 1e-14 * s
    go.deconv()
    gg.deconv()
    assert np.abs(gg.fk() - go.fk()).max() <= 1e-13 * np.abs(go.fk()).max()


@pytest.mark.parametrize("Ng,Np,interlace", [(32, 20000, True), (48, 30000, False), (64, 50000, True),
                                             (45, 30000, True), (34, 20000, False),
                                             (37, 20000, True)])
def test_pipeline_bins_vs_oracle(torch_cuda, lib, orc, Ng, Np, interlace):
    """fft_p2k (or p2g+x2k+deconv) -> P -> Pl/Pmu bins, reference call order
    (src/app/P.c:53-66); counts exact, sums to FP64 rounding.  Odd sides and
    sides with a prime factor above 13 included."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    L = 500.0
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=Ng, Ngs=32, dtype=np.float64)
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    if interlace:
        go.p2k(co)
        gg.p2k(cg)
    else:
        for g, c in ((go, co), (gg, cg)):
            g.p2g(c, (0, 0, 0))
            g.x2k(0)
            g.deconv()
    go.P(co)
    gg.P(cg)
    dK = 2 * np.pi / L
    for los in ((0, 0, 1), (1, 1, 0), (0, 0, 0)):
        Ko, No, Po = go.Pl_bins(los, dK)
        Kg, Ng_, Pg = gg.Pl_bins(los, dK)
        assert np.array_equal(No, Ng_)
        ok = No > 0
        assert np.allclose(Kg[ok], Ko[ok], rtol=1e-13, atol=0)
        assert np.all(np.isnan(Kg[~ok]))
        scale = np.abs(Po[ok, 0]).max()
        assert np.abs(Pg[ok] - Po[ok]).max() <= 1e-10 * scale
        for Nmu in (3, 10):
            Ko, No, Po = go.Pmu_bins(los, dK, Nmu)
            Kg, Ng_, Pg = gg.Pmu_bins(los, dK, Nmu)
            assert np.array_equal(No, Ng_)
            assert np.allclose(Kg, Ko, rtol=1e-13, atol=0)
            assert np.abs(Pg - Po).max() <= 1e-10 * np.abs(Po).max()


def _read_table(path):
    head, rows = [], []
    with open(path) as fp:
        for line in fp:
            (head if line.startswith("#") else rows).append(line)
    return head, np.array([[float(v) for v in r.split()] for r in rows])


def test_text_files_vs_oracle(torch_cuda, lib, orc, tmp_path):
    """Pl/Pmu files: identical headers and counts, values to FP64 rounding
    (src/power.c:78-99, 155-176)."""
    import fftgal_amd as fa
    rng = np.random.default_rng(11)
    Ng, L = 16, 200.0
    x, y, z = rng.random((3, 4000)) * L
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    go.p2k(co)
    gg.p2k(cg)
    go.P(co)
    gg.P(cg)
    dK = 2 * np.pi / L
    for kind in ("Pl", "Pmu"):
        po, pg = str(tmp_path / f"{kind}_o.txt"), str(tmp_path / f"{kind}_g.txt")
        lo = getattr(go, kind)(co, (1, 0, 1), dK, po)
        lg = getattr(gg, kind)(cg, (1, 0, 1), dK, pg)
        assert np.array_equal(lo, lg)   # hat() normalised the caller's los
        ho, to = _read_table(po)
        hg, tg = _read_table(pg)
        assert ho == hg
        assert to.shape == tg.shape
        assert np.array_equal(np.isnan(to), np.isnan(tg))
        m = ~np.isnan(to)
        assert np.allclose(tg[m], to[m], rtol=2e-6, atol=1e-300)   # %e prints 7 digits


def test_resident_mode_and_export(torch_cuda, lib, orc):
    import fftgal_amd as fa
    rng = np.random.default_rng(5)
    Ng, L = 24, 300.0
    x, y, z = rng.random((3, 5000)) * L
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    go = orc.Grid(Ng, L)
    gg = fa.Grid(Ng, L, resident=True)
    go.p2k(co)
    gg.p2k(cg)
    assert np.abs(gg.f).max() == 0.0          # host mirror untouched in resident mode
    gg.sync_host()
    assert np.abs(gg.fk() - go.fk()).max() <= 1e-13 * np.abs(go.fk()).max()
    exp = gg.lib.fft_exportf(gg.ptr)
    arr = np.ctypeslib.as_array(exp, shape=(gg.ptr.contents.Ng3_pad,)).copy()
    assert np.array_equal(arr.reshape(gg.f.shape), gg.f)
    import ctypes
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(exp, ctypes.c_void_p))


def test_c_octet_program(torch_cuda, lib):
    """The octet test as a C program compiled against include/ and linked
    with libfftgal_amd.so (tests/c, built by __graft_entry__.build())."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "test_octet")
    assert os.path.exists(exe), "tests/c/test_octet not built"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "test PASSED" in out.stdout


class _Ptr:
    def __init__(self, p):
        self.ptr = p


def test_P_driver_flow_vs_oracle(torch_cuda, lib, orc, tmp_path):
    """The reference's P.c flow (src/app/P.c:19-97) as a C program on the
    unchanged API (tests/c/test_pflow.c: QPM load, six lines of sight in
    redshift space, then real space reusing one transform through
    fft_exportf / fft_importf), against the same sequence on the oracle:
    all 24 text files with identical headers and counts, values to the 7
    printed digits."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "test_pflow")
    assert os.path.exists(exe), "tests/c/test_pflow not built"
    rng = np.random.default_rng(17)
    n, L, Ng, fold, dK, a = 20000, 500.0, 32, 2, 0.02, 0.6
    pos = rng.random((n, 3)) * L
    vel = rng.standard_normal((n, 3)) * 400.0
    path = str(tmp_path / "mock.qpm")
    with open(path, "w") as fp:
        for p, v in zip(pos, vel):
            fp.write(f"{p[0]:.5f} {p[1]:.5f} {p[2]:.5f} {v[0]:.3f} {v[1]:.3f} {v[2]:.3f} "
                     f"2.5e+12 1 7\n")
    gdir, odir = tmp_path / "gpu", tmp_path / "orc"
    gdir.mkdir()
    odir.mkdir()
    out = subprocess.run([exe, str(Ng), str(L), str(fold), str(dK), path, str(a), str(gdir)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    # the same flow on the oracle
    o = orc.load()
    part = _Ptr(o.oracle_loadqpm_cubic(path.encode(), L))
    g = orc.Grid(Ng, L / fold)
    aH = a * 100 * np.sqrt(0.29 / a ** 3 + 0.71)
    los_all = [(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)
               if 0 < i + j + k < 3]
    # one los array per line of sight through gal_rsd, Pl and Pmu, each of
    # which normalises it in place (src/gal.c:64, src/power.c:33,107), as
    # P.c passes it: a second normalisation can change its last bit, and
    # with it the wedge of modes with |mu| = 1/3 or 2/3 exactly
    dp = C.POINTER(C.c_double)
    for los in los_all:
        lv = np.array(los, dtype=np.float64)
        rsd = _Ptr(o.oracle_gal_rsd(part.ptr, lv.ctypes.data_as(dp), aH))
        g.p2k(rsd)
        g.P(rsd)
        tag = "%d%d%d" % los
        o.oracle_Pl(g.ptr, rsd.ptr, lv.ctypes.data_as(dp), dK,
                    str(odir / f"Pl_rsd1_los{tag}_fd{fold}.txt").encode())
        o.oracle_Pmu(g.ptr, rsd.ptr, lv.ctypes.data_as(dp), dK,
                     str(odir / f"Pmu_rsd1_los{tag}_fd{fold}.txt").encode())
        o.oracle_gal_free(rsd.ptr)
    g.p2k(part)
    fk = o.oracle_fft_exportf(g.ptr)
    for los in los_all:
        o.oracle_fft_importf(g.ptr, fk)
        g.P(part)
        tag = "%d%d%d" % los
        lv = np.array(los, dtype=np.float64)
        o.oracle_Pl(g.ptr, part.ptr, lv.ctypes.data_as(dp), dK,
                    str(odir / f"Pl_rsd0_los{tag}_fd{fold}.txt").encode())
        o.oracle_Pmu(g.ptr, part.ptr, lv.ctypes.data_as(dp), dK,
                     str(odir / f"Pmu_rsd0_los{tag}_fd{fold}.txt").encode())
    import ctypes
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(fk, ctypes.c_void_p))
    o.oracle_gal_free(part.ptr)
    names = sorted(os.listdir(odir))
    assert len(names) == 24 and sorted(os.listdir(gdir)) == names
    for name in names:
        ho, to = _read_table(str(odir / name))
        hg, tg = _read_table(str(gdir / name))
        assert ho == hg, name
        assert to.shape == tg.shape, name
        assert np.array_equal(np.isnan(to), np.isnan(tg)), name
        m = ~np.isnan(to)
        # 7 printed digits; values that cancel to near zero (a wedge's P
        # after the shot-noise subtraction) differ by the FP64 summation
        # order, far below the column's scale
        scale = np.abs(to[m]).max()
        bad = ~np.isclose(tg[m], to[m], rtol=2e-6, atol=1e-12 * scale)
        assert not bad.any(), (name, tg[m][bad][:5], to[m][bad][:5])


@pytest.mark.parametrize("kind,Nsub", [(0, 2), (0, 4), (1, 2), (1, 4)])
def test_dl_filter_vs_oracle(torch_cuda, lib, orc, kind, Nsub):
    """fft_dl_filter (the super-sample filter of src/app/DLsb.c:66-96 and
    DLss.c:62-89 on the device) against the oracle's restatement: painted
    catalog as DLsb paints it (offset 1/2 for the sub box, 0 for the sphere),
    fft_x2k without phase, then per line of sight and ell the filter from the
    same restored spectrum; corners and the whole configuration-space grid
    to FP64 rounding."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    Ng, L, Np = 32, 400.0, 30000
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=9, Ngs=32, dtype=np.float64)
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    off = 0.5 if kind == 0 else 0.0
    go.p2g(co, (off, off, off))
    gg.p2g(cg, (off, off, off))
    go.x2k(0)
    gg.x2k(0)
    fk_o, fk_g = go.f.copy(), gg.f.copy()
    dp = C.POINTER(C.c_double)
    for los in ((0.0, 0.0, 1.0), (2 ** -0.5, 2 ** -0.5, 0.0)):
        lv = np.array(los)
        for ell in (0, 2, 4):
            go.f[:] = fk_o
            gg.f[:] = fk_g
            do = np.zeros(Nsub ** 3)
            dg = np.zeros(Nsub ** 3)
            o = orc.load()
            o.oracle_dl_filter(go.ptr, kind, Nsub, lv.ctypes.data_as(dp), ell, 1.0,
                               do.ctypes.data_as(dp))
            lib.fft_dl_filter(gg.ptr, kind, Nsub, lv.ctypes.data_as(dp), ell, 1.0,
                              dg.ctypes.data_as(dp))
            sc = np.abs(do).max()
            assert np.abs(dg - do).max() <= 1e-12 * sc, (los, ell)
            assert np.abs(gg.f[:, :, :Ng] - go.f[:, :, :Ng]).max() <= \
                1e-12 * np.abs(go.f[:, :, :Ng]).max()


@pytest.mark.parametrize("kind,Nsub", [(0, 2), (1, 2), (1, 4)])
def test_subvolume_flow_vs_oracle(torch_cuda, lib, orc, kind, Nsub):
    """The sub-volume flow of src/app/Psb.c:54-96 and Pss.c:57-96 for one
    sub-volume: gal_subbox / gal_subsphere with randoms, fft_p2k, P (with
    alpha and the randoms' shot noise), the window deconvolution on the
    device (fft_subvolume_deconv), Pl / Pmu bins; against the oracle:
    identical catalogs, counts exact, sums to FP64 rounding, and the whole
    deconvolved grid."""
    import fftgal_amd as fa
    Ng, L = 32, 200.0
    rng = np.random.default_rng(23)
    x, y, z = rng.random((3, 30000)) * L
    co, cg = _catalogs(orc, fa, x, y, z, L ** 3)
    dp = C.POINTER(C.c_double)
    gs = L / Nsub
    if kind == 0:
        v = np.array([gs, 2 * gs, 0.0, gs, gs, 2 * gs, L])
        so = _Ptr(orc.load().oracle_gal_subbox(co.ptr, v.ctypes.data_as(dp), 0.05))
        sg = _Ptr(lib.gal_subbox(cg.ptr, v.ctypes.data_as(dp), 0.05))
    else:
        v = np.array([gs, 0.0, gs, gs * np.sqrt(3) / 2, L])
        so = _Ptr(orc.load().oracle_gal_subsphere(co.ptr, v.ctypes.data_as(dp), 0.05))
        sg = _Ptr(lib.gal_subsphere(cg.ptr, v.ctypes.data_as(dp), 0.05))
    a, b = so.ptr.contents, sg.ptr.contents
    assert a.Np == b.Np and a.rand.contents.Np == b.rand.contents.Np
    n, nr = b.Np, b.rand.contents.Np
    assert np.array_equal(np.ctypeslib.as_array(a.x, (n,)), np.ctypeslib.as_array(b.x, (n,)))
    assert np.array_equal(np.ctypeslib.as_array(a.rand.contents.z, (nr,)),
                          np.ctypeslib.as_array(b.rand.contents.z, (nr,)))
    go, gg = orc.Grid(Ng, L), fa.Grid(Ng, L)
    go.p2k(so)
    gg.p2k(sg)
    go.P(so)
    gg.P(sg)
    orc.load().oracle_subvolume_deconv(go.ptr, kind, Nsub)
    lib.fft_subvolume_deconv(gg.ptr, kind, Nsub)
    assert np.abs(gg.f - go.f).max() <= 1e-10 * np.abs(go.f).max()
    dK = 2 * np.pi / L
    for los in ((0, 0, 1), (0, 1, 1)):
        Ko, No, Po = go.Pl_bins(los, dK)
        Kg, Ng_, Pg = gg.Pl_bins(los, dK)
        assert np.array_equal(No, Ng_)
        ok = No > 0
        assert np.abs(Pg[ok] - Po[ok]).max() <= 1e-9 * np.abs(Po[ok, 0]).max()
        Ko, No, Po = go.Pmu_bins(los, dK, 3)
        Kg, Ng_, Pg = gg.Pmu_bins(los, dK, 3)
        assert np.array_equal(No, Ng_)
        assert np.abs(Pg - Po).max() <= 1e-9 * np.abs(Po).max()
    orc.load().oracle_gal_free(so.ptr)
    lib.gal_free(sg.ptr)


def test_fft_init_wisdom_measure_export_import(torch_cuda, lib, orc, tmp_path):
    """fft_init honours `wisdom` like src/fft.c:30-44: a file without a line
    for this grid is measured and written, a second call imports the line,
    and the transform is the same FP64 FFT either way."""
    import fftgal_amd as fa
    Ng, L = 48, 100.0
    w = tmp_path / "48.wsdm"
    rng = np.random.default_rng(5)
    x = rng.standard_normal((Ng, Ng, Ng))
    ref = orc.Grid(Ng, L)
    ref.f[:, :, :Ng] = x
    ref.x2k(0)
    scale = np.abs(ref.fk()).max()
    for k in range(2):
        g = fa.Grid(Ng, L, wisdom=str(w))
        g.f[:, :, :Ng] = x
        g.x2k(0)
        assert np.abs(g.fk() - ref.fk()).max() <= 1e-13 * scale
        lines = [ln.split() for ln in w.read_text().splitlines()
                 if ln.strip() and not ln.startswith("#")]
        assert len(lines) == 1 and lines[0][:2] == ["48", "2"], lines
        g.free()
    # a line for another grid side leaves this one to be measured and appended
    fa.Grid(30, L, wisdom=str(w)).free()
    sides = sorted(int(ln.split()[0]) for ln in w.read_text().splitlines()
                   if ln.strip() and not ln.startswith("#"))
    assert sides == [30, 48]


"""GPU parity of the mixed-precision pipeline (include/fftgal_batch.h)
against the FP64 oracle: deterministic fixed-point paint, FP32 FFT, fused
bin kernel.  Mode counts bit-exact; P_l within a stated tolerance."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(torch, *arrs):
    ts = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda() for a in arrs]
    torch.cuda.synchronize()
    return ts


def _oracle_bins(orc, x, y, z, V, Ng, L, interlace, los, dK, Nmu):
    return orc.power_pipeline(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64),
                              V, Ng, L, interlace=interlace, los=los, dK=dK, Nmu=Nmu)


def _rel_err(a, b, ref_scale):
    return np.abs(a - b) / ref_scale


def _paint_vs_oracle(torch, orc, x, y, z, L, Ng, tile):
    """Fixed-point paint of float32 positions against the oracle's FP64
    fft_p2g at offsets 0 and 1/2, within the stated per-cell bound; mass
    deficit and run-to-run bit identity."""
    import fftgal_amd as fa
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, Np_max=len(x), paint_tile=tile)
    A, B = p.paint_only(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), len(x))
    # each footprint rounds to 2^-S particle or finer: S >= 18 (a pencil
    # segment's two neighbouring tiles of up to 8192 particles each), 19 for a
    # full 8192-particle dense item
    S = p.fixed_bits(16384)
    assert S == 18 and p.fixed_bits(8192) == 19
    assert p.fixed_bits(1) == 25 and p.fixed_bits(216) == 25
    unit_fx = p.fixed_scale()
    assert abs(unit_fx * 2.0 ** 25 - 1) <= 1.1e-4
    a = A.astype(np.float64) * unit_fx
    b = B.astype(np.float64) * unit_fx
    # oracle grids in particle units: undo 1/(nbar H^3)
    V = L ** 3
    H3 = (L / Ng) ** 3
    unit = 1.0 / (len(x) / V) / H3
    c = orc.Catalog(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64), V)
    g = orc.Grid(Ng, L)
    g.p2g(c, (0, 0, 0))
    ra = g.f[:, :, :Ng] / unit
    g.p2g(c, (0.5, 0.5, 0.5))
    rb = g.f[:, :, :Ng] / unit
    # Stated tolerance per stencil term: dithered weight rounding (< 2^-S),
    # the packed tile-local position (F = min(23 - log2(T), 24) fraction
    # bits, so at most 2^-(F+1) cell away, moving a weight product by
    # <= 0.67 * 2^-(F+1)), and FP32 weight evaluation (1e-7).  The oracle
    # sees the same float32 positions.  A cell's bound is that times the
    # number of particles whose 4x4x4 stencil covers it (counted here with
    # numpy).
    T = min(tile, Ng)
    F = min(23 - int(np.ceil(np.log2(T))), 24)
    term = 2.0 ** -S + 0.34 * 2.0 ** -F + 1e-7
    for grid, ref, off in ((a, ra, 0.0), (b, rb, 0.5)):
        cnt = np.zeros((Ng, Ng, Ng))
        base = [np.floor(v.astype(np.float64) * Ng / L - off).astype(np.int64)
                for v in (x, y, z)]
        for di in range(-1, 3):
            for dj in range(-1, 3):
                for dk in range(-1, 3):
                    np.add.at(cnt, ((base[0] + di) % Ng, (base[1] + dj) % Ng,
                                    (base[2] + dk) % Ng), 1.0)
        assert np.all(np.abs(grid - ref) <= cnt * term + 1e-12)
    # mass: every particle contributes 1 up to 64 rounding steps of less
    # than 2^-S each.  The rounding is dithered (unbiased), so the total
    # error is a zero-mean sum over particles: its standard deviation is at
    # most sqrt(N) * 64/sqrt(12) * 2^-S even if a particle's 64 errors were
    # fully correlated; 6 sigma of that (a deterministic check for this seed)
    for grid in (a, b):
        deficit = len(x) - grid.sum()
        assert abs(deficit) <= len(x) * 64 * 2.0 ** -S
        assert abs(deficit) <= 6 * np.sqrt(len(x)) * 64 / np.sqrt(12) * 2.0 ** -S
    # determinism: a second paint gives bit-identical integer grids
    A2, B2 = p.paint_only(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), len(x))
    assert np.array_equal(A, A2) and np.array_equal(B, B2)
    p.close()


@pytest.mark.parametrize("Ng,tile", [(32, 16), (48, 16), (64, 8), (12, 16), (48, 1), (40, 16)])
def test_fixed_point_paint_vs_oracle(torch_cuda, lib, orc, Ng, tile):
    """(40, 16): the tiles do not divide the grid, so the last tile per axis
    wraps onto the first."""
    from fftgal_amd import mockgen
    L = 400.0
    (x, y, z), _ = mockgen.lognormal(40000, L, seed=Ng, Ngs=32)
    # positions also outside [0, L): the paint wraps them
    x = x - 0.3 * L
    _paint_vs_oracle(torch_cuda, orc, x, y, z, L, Ng, tile)


def test_single_cell_catalog_paint_vs_oracle(torch_cuda, lib, orc):
    """Every particle in one grid cell: each wave of both sort levels holds
    a single digit (the one-atomic path of the counts), and the one tile
    takes several work items."""
    rng = np.random.default_rng(8)
    L, Ng = 400.0, 64
    H = L / Ng
    pos = rng.uniform(37 * H, 38 * H, size=(20000, 3)).astype(np.float32)
    _paint_vs_oracle(torch_cuda, orc, pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), L,
                     Ng, 16)


def test_dense_tile_paint_vs_oracle(torch_cuda, lib, orc):
    """One tile holds 30000 particles (more than one work item's 8192), the
    rest of the box a thin uniform field: tiles painted by several work
    items next to tiles painted by one."""
    rng = np.random.default_rng(5)
    L, Ng = 400.0, 64
    H = L / Ng
    clump = rng.uniform(20 * H, 23 * H, size=(30000, 3))
    field = rng.uniform(0.0, L, size=(20000, 3))
    pos = np.concatenate([clump, field]).astype(np.float32)
    pos = pos[rng.permutation(len(pos))]
    _paint_vs_oracle(torch_cuda, orc, pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), L,
                     Ng, 16)


def _unpack_spectrum(c):
    """packed half spectrum [Ng, Ng, Ng/2] -> r2c layout [Ng, Ng, Ng/2+1]:
    column 0 holds Z = F(X0) + i F(XN)"""
    Ng = c.shape[0]
    c = c.astype(np.complex128)
    m = (-np.arange(Ng)) % Ng
    Zm = np.conj(c[m][:, m, 0])
    out = np.empty((Ng, Ng, Ng // 2 + 1), np.complex128)
    out[:, :, 1:Ng // 2] = c[:, :, 1:]
    out[:, :, 0] = (c[:, :, 0] + Zm) / 2
    out[:, :, Ng // 2] = (c[:, :, 0] - Zm) / 2j
    return out


@pytest.mark.parametrize("Ng,generic", [(32, False), (48, False), (64, False), (96, False),
                                        (64, True), (40, False), (80, False)])
def test_fp32_fft_vs_numpy(torch_cuda, lib, Ng, generic):
    """The FP32 spectra of the fast path (specialised kernels, or the generic
    ones for Ng = 40, 80 and when forced) against numpy's FP64 rfftn of the
    same fixed-point grids."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    L, Np = 300.0, 60000
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=Ng, Ngs=32)
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, Nmu=4, Np_max=Np)
    used = p.fft_generic(generic)
    assert used == (generic or Ng in (40, 80))
    A, B = p.paint_only(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np)
    p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, L ** 3)
    sA, sB = p.spectra_host()
    for grid, spec in ((A, sA), (B, sB)):
        ref = np.fft.rfftn(grid.astype(np.float64), axes=(0, 1, 2))
        got = _unpack_spectrum(spec)
        # FP32 transform: relative L2 error a few FP32 ulps times log2 of the size
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel < 2e-6, rel
        assert np.abs(got - ref).max() <= 5e-6 * np.abs(ref).max()
    p.close()


@pytest.mark.parametrize("Ng,Np,interlace,Nmu", [(32, 30000, True, 10), (64, 200000, True, 10),
                                                 (48, 100000, False, 3), (96, 300000, True, 20),
                                                 (40, 80000, True, 5)])
def test_fast_pipeline_vs_oracle(torch_cuda, lib, orc, Ng, Np, interlace, Nmu):
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    L = 600.0
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=Ng + Np, Ngs=64)
    V = L ** 3
    los = (0.0, 0.0, 1.0)
    dK = 2 * np.pi / L
    ref = _oracle_bins(orc, x, y, z, V, Ng, L, interlace, los, dK, Nmu)
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=interlace, dK=dK, los=los, Nmu=Nmu, Np_max=Np)
    out = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, V)
    # bit-exact mode counts, Legendre and wedges
    assert np.array_equal(out["N"], ref["N"])
    assert np.array_equal(out["Nmu"], ref["Nmu"])
    ok = ref["N"] > 0
    assert np.allclose(out["K"][ok], ref["K"][ok], rtol=1e-12, atol=0)
    # P_l: |dP_l| <= 1e-5 (|P_l| + P_0 + Pshot)... relative to the bin's power
    kmax = Ng // 2            # k < k_Nyquist
    sel = ok & (np.arange(len(ok)) < kmax)
    Pshot = V / Np
    scale = np.abs(ref["Pl"][sel, 0]) + Pshot
    for l in range(3):
        err = np.abs(out["Pl"][sel, l] - ref["Pl"][sel, l]) / scale
        assert err.max() <= 1e-5, (l, err.max())
    wm = ref["Nmu"][:kmax] > 0
    errw = np.abs(out["Pmu"][:kmax][wm] - ref["Pmu"][:kmax][wm]) / (
        np.abs(ref["Pmu"][:kmax][wm]) + Pshot)
    assert errw.max() <= 1e-5
    p.close()


@pytest.mark.parametrize("Np", [1, 63])
def test_tiny_catalogs_vs_oracle(torch_cuda, lib, orc, Np):
    """One particle, and fewer than a wave: the sort's partial chunks and
    waves, one work item, against the oracle pipeline."""
    import fftgal_amd as fa
    torch = torch_cuda
    rng = np.random.default_rng(Np)
    Ng, L = 32, 300.0
    x, y, z = (rng.random((3, Np)) * L).astype(np.float32)
    V = L ** 3
    dK = 2 * np.pi / L
    ref = _oracle_bins(orc, x, y, z, V, Ng, L, True, (0, 0, 1), dK, 4)
    t = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, dK=dK, los=(0, 0, 1), Nmu=4, Np_max=Np)
    out = p.run(*[a.data_ptr() for a in t], Np, V)
    p.close()
    assert np.array_equal(out["N"], ref["N"])
    ok = (ref["N"] > 0) & (np.arange(len(ref["N"])) < Ng // 2)
    scale = np.abs(ref["Pl"][ok, 0]) + V / Np
    for l in range(3):
        assert (np.abs(out["Pl"][ok, l] - ref["Pl"][ok, l]) / scale).max() <= 1e-5


@pytest.mark.parametrize("Ng,n1", [(256, 16), (384, 16)])
def test_fourstep_i_pass_matches_one_pass(torch_cuda, lib, Ng, n1):
    """The i-axis column pass as a four-step n1 x Ng/n1 (what grids of
    Ng >= 2048 run, forced here on small grids): frequency i lands on plane
    (i % n1) * (Ng / n1) + i / n1 of the spectra, equal to the one-pass
    transform to FP32 rounding, and the bins read it through that map."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    L = 800.0
    (x, y, z), _ = mockgen.lognormal(300000, L, seed=Ng, Ngs=32)
    t = _dev(torch, x, y, z)
    ptr = [a.data_ptr() for a in t]
    V = L ** 3
    res, spec = {}, {}
    for split in (0, n1):
        os.environ["FFTGAL_FFT_FOURSTEP"] = str(split)
        try:
            p = fa.PowerPlan(Ng, L, interlace=True, Nmu=6, Np_max=len(x))
        finally:
            os.environ.pop("FFTGAL_FFT_FOURSTEP", None)
        res[split] = p.run(*ptr, len(x), V)
        if Ng <= 256:
            spec[split] = p.spectra_host()[0]
        p.close()
    if spec:
        i = np.arange(Ng)
        plane = (i % n1) * (Ng // n1) + i // n1
        ref = spec[0]
        assert np.abs(spec[n1][plane] - ref).max() <= 1e-5 * np.abs(ref).max()
    a, b = res[0], res[n1]
    assert np.array_equal(a["N"], b["N"])
    ok = a["N"] > 0
    scale = np.abs(a["Pl"][ok, 0]) + V / len(x)
    for l in range(4):
        assert (np.abs(b["Pl"][ok, l] - a["Pl"][ok, l]) / scale).max() <= 1e-5
    wm = a["Nmu"] > 0
    assert np.array_equal(a["Nmu"], b["Nmu"])
    assert (np.abs(b["Pmu"][wm] - a["Pmu"][wm]) / (np.abs(a["Pmu"][wm]) + V / len(x))).max() <= 1e-5


def test_graph_replays_match_eager(torch_cuda, lib):
    """The captured graph replayed several times, with eager work between
    replays, gives the same bins as plain stream launches."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    Ng, L, Np = 64, 500.0, 150000
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=5, Ngs=64)
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, Nmu=10, Np_max=Np)
    p.use_graph(False)
    eager = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, L ** 3)
    p.use_graph(True)
    # integer grids and exact fixed-point bin sums: bit for bit
    for r in range(4):
        out = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, L ** 3)
        for key in ("N", "Nmu", "K", "Kmu", "Pl", "Pmu"):
            assert np.array_equal(out[key], eager[key], equal_nan=True), (r, key)
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        del junk
    p.close()


@pytest.mark.parametrize("sets", [2, 1])
def test_consecutive_catalogs_match_fresh_plans(torch_cuda, lib, monkeypatch, sets):
    """Catalogs run back to back through one plan, with two int64 grid sets
    (catalogs alternate, each paint clears the other set) or one (the row
    FFT clears it): graph and eager runs, after paint_only and after a
    profile, each give the bins of a fresh plan for that catalog."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    Ng, L = 64, 500.0
    cats = {}
    for name, (Np, seed) in {"A": (150000, 5), "B": (90000, 6)}.items():
        (x, y, z), _ = mockgen.lognormal(Np, L, seed=seed, Ngs=64)
        cats[name] = (_dev(torch, x, y, z), Np)
    V = L ** 3
    ref = {}
    for name, (t, Np) in cats.items():
        q = fa.PowerPlan(Ng, L, interlace=True, Nmu=5, Np_max=Np)
        ref[name] = q.run(*[a.data_ptr() for a in t], Np, V)
        q.close()

    def check(out, name):
        r = ref[name]
        tol = 1e-12 * np.nanmax(np.abs(r["Pl"][:, 0]))
        assert np.array_equal(out["N"], r["N"]) and np.array_equal(out["Nmu"], r["Nmu"])
        assert np.allclose(out["Pl"], r["Pl"], rtol=1e-12, atol=tol, equal_nan=True), name
        assert np.allclose(out["Pmu"], r["Pmu"], rtol=1e-12, atol=tol, equal_nan=True), name

    if sets == 1:
        monkeypatch.setenv("FFTGAL_GRID_SETS", "1")
    p = fa.PowerPlan(Ng, L, interlace=True, Nmu=5, Np_max=150000)
    monkeypatch.delenv("FFTGAL_GRID_SETS", raising=False)
    assert p.grid_sets() == sets
    for step in ("A", "B", "B", "A", "paint_only A", "B", "profile A", "A", "eager B",
                 "eager A", "A", "B"):
        name = step.split()[-1]
        t, Np = cats[name]
        ptr = [a.data_ptr() for a in t]
        if step.startswith("paint_only"):
            p.paint_only(*ptr, Np)
            continue
        if step.startswith("profile"):
            p.profile(*ptr, Np, V)
            continue
        p.use_graph(not step.startswith("eager"))
        check(p.run(*ptr, Np, V), name)
    p.close()


@pytest.mark.parametrize("Ng,tile", [(96, 2), (160, 1)])
def test_large_tile_count_sorts(torch_cuda, lib, orc, Ng, tile):
    """Grids with more paint tiles than 256 x 128: 48^3 tiles of side 2 at
    Ng = 96 take the wide two-level sort (up to 2048 x 1024 digits), 160^3
    tiles of side 1 at Ng = 160 (over 2^21) the single-level sort; bins
    against the FP64 oracle."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    Np, L = 200000, 600.0
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=11, Ngs=64)
    V = L ** 3
    dK = 2 * np.pi / L
    ref = _oracle_bins(orc, x, y, z, V, Ng, L, True, (0, 0, 1), dK, 5)
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, dK=dK, los=(0, 0, 1), Nmu=5, Np_max=Np,
                     paint_tile=tile)
    assert p.tuning()["paint_tile"] == tile
    out = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, V)
    assert np.array_equal(out["N"], ref["N"])
    assert np.array_equal(out["Nmu"], ref["Nmu"])
    sel = (ref["N"] > 0) & (np.arange(len(ref["N"])) < Ng // 2)
    scale = np.abs(ref["Pl"][sel, 0]) + V / Np
    for l in range(3):
        assert (np.abs(out["Pl"][sel, l] - ref["Pl"][sel, l]) / scale).max() <= 1e-5
    p.close()


def test_ng1024_fast_vs_plain_fp64(torch_cuda, lib):
    """Ng = 1024 (specialised length-1024 columns, length-512 rows, 64^3 tiles):
    the fast path against the plain FP64 GPU form of the drop-in API on the
    same 4,000,000-particle clustered catalog; counts exact, P_l within the
    stated tolerance below k_Nyquist."""
    import fftgal_amd as fa
    from fftgal_amd import mockgen
    torch = torch_cuda
    Ng, Np, L = 1024, 4_000_000, 2000.0
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=3, Ngs=256)
    V = L ** 3
    dK = 2 * np.pi / L
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, dK=dK, los=(0, 0, 1), Nmu=4, Np_max=Np)
    out = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, V)
    p.close()
    del tx, ty, tz
    torch.cuda.empty_cache()
    dd = [torch.from_numpy(a.astype(np.float64)).cuda() for a in (x, y, z)]
    torch.cuda.synchronize()
    g = fa.Grid(Ng, L, resident=True)
    g.lib.fft_p2k_device(g.ptr, dd[0].data_ptr(), dd[1].data_ptr(), dd[2].data_ptr(), Np, V)
    g.lib.P_params(g.ptr, Np, V, 0)
    K, N, Pl = g.Pl_bins((0, 0, 1), dK)
    Km, Nm, Pm = g.Pmu_bins((0, 0, 1), dK, 4)
    g.free()
    assert np.array_equal(out["N"], N)
    assert np.array_equal(out["Nmu"], Nm)
    sel = (N > 0) & (np.arange(len(N)) < Ng // 2)
    scale = np.abs(Pl[sel, 0]) + V / Np
    for l in range(3):
        err = (np.abs(out["Pl"][sel, l] - Pl[sel, l]) / scale).max()
        print(f"l={2 * l}: max |dP|/(|P0|+Pshot) = {err:.2e}")
        assert err <= 1e-5


@pytest.mark.parametrize("Ng", [256, 2048])
def test_lattice_shot_noise_fast_and_plain(torch_cuda, lib, Ng):
    """Closed form at up to the largest grid (Ng = 2048, 68.7 GB per FP64
    grid): particles on a cubic sublattice of spacing s = 8 cells (offset
    0.3 cell) make a density field of period s cells, whose Fourier modes
    vanish except at multiples of Ng/s.  Below that wavenumber every mode's
    power is -Pshot, and since the shot noise is subtracted per mode before
    the Legendre weighting (src/power.c:18-24, 59-67), bin b holds
    P_l = -Pshot (2l+1) <L_l(mu)> over the integer lattice vectors of the
    shell b <= |n| < b+1 (numpy enumeration, first 30 shells; the rest:
    fast path against the FP64 form).  Both forms must give that, and the
    same mode counts."""
    import fftgal_amd as fa
    torch = torch_cuda
    L, s = 4000.0, 8
    H = L / Ng
    n = Ng // s
    g1 = (np.arange(n) * s + 0.3) * H
    x, y, z = (a.ravel().astype(np.float32) for a in np.meshgrid(g1, g1, g1, indexing="ij"))
    Np = x.size
    V = L ** 3
    dK = 2 * np.pi / L
    Pshot = V / Np
    kmax = n - 2
    tx, ty, tz = _dev(torch, x, y, z)
    p = fa.PowerPlan(Ng, L, interlace=True, dK=dK, los=(0, 0, 1), Nmu=4, Np_max=Np)
    out = p.run(tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), Np, V)
    p.close()
    torch.cuda.empty_cache()
    dd = [t.double() for t in (tx, ty, tz)]
    del tx, ty, tz
    torch.cuda.synchronize()
    g = fa.Grid(Ng, L, resident=True)
    g.lib.fft_p2k_device(g.ptr, dd[0].data_ptr(), dd[1].data_ptr(), dd[2].data_ptr(), Np, V)
    g.lib.P_params(g.ptr, Np, V, 0)
    K, N, Pl = g.Pl_bins((0, 0, 1), dK)
    g.free()
    del dd
    torch.cuda.empty_cache()
    assert np.array_equal(out["N"], N)
    # closed form of the first shells
    nb = 30
    r = np.arange(-nb, nb + 1, dtype=np.float64)
    fi, fj, fk = np.meshgrid(r, r, r, indexing="ij")
    kr = np.sqrt(fi * fi + fj * fj + fk * fk).ravel()
    mu = np.divide(fk.ravel(), kr, out=np.zeros_like(kr), where=kr > 0)
    ib = np.floor(kr).astype(np.int64)
    use = (kr > 0) & (ib < nb)
    cnt = np.bincount(ib[use], minlength=nb)
    m2 = mu[use] ** 2
    leg = (np.ones_like(m2), 1.5 * m2 - 0.5, (4.375 * m2 - 3.75) * m2 + 0.375)
    exact = np.stack([-Pshot * (4 * l + 1) * np.bincount(ib[use], weights=leg[l],
                                                         minlength=nb) / np.maximum(cnt, 1)
                      for l in range(3)], axis=1)
    assert np.array_equal(N[1:nb], cnt[1:nb])
    # the fast path's FP32 FFT leaks O(eps) amplitude into the empty modes,
    # O(eps^2) in power, and its bin sums are fixed point: each value (at
    # most 4 modes) rounds to 2^-24 of Pshot's power of two, below 2^-25
    # Pshot: 2^-24 Pshot on a bin; the FP64 form O(eps_64^2)
    for Pv, tol in ((out["Pl"], 2.0 ** -24), (Pl, 1e-12)):
        assert (np.abs(Pv[1:nb, :3] - exact[1:nb]) / Pshot).max() <= tol
        assert (np.abs(Pv[1:kmax, 0] + Pshot) / Pshot).max() <= tol
    assert (np.abs(out["Pl"][1:kmax, :3] - Pl[1:kmax, :3]) / Pshot).max() <= 2.0 ** -24

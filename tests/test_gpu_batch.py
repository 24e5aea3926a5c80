"""GPU parity of batches of catalogs (fftgal_batch_*, include/fftgal_batch.h):
many catalogs on grids of different sides, boxes (folding) and interlacing,
sorted, painted, transformed and binned with one launch per stage.

Each catalog of a batch must give what a plan of its own gives
(fftgal_power_t, pinned against the FP64 oracle in test_gpu_fast.py): the
int64 grids and the FP32 spectra are the same computations and the bin
sums exact fixed point, so mode counts are equal and P_l and the wedges
differ at most by the fixed-point rounding of sums grouped differently
(stated tolerance 1e-12 of |P_0| + Pshot).  One folded catalog is also
checked against the oracle directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[1, 0], ids=["one_launch", "per_side"])
def _launch_form(request, monkeypatch):
    """Every test twice: rounds of several grid sides run each FFT pass,
    the paint and the fused bin as one launch (multi-size kernels), or one
    launch per grid or tile side (FFTGAL_BATCH_MULTI=0); both must give the
    same bins."""
    monkeypatch.setenv("FFTGAL_BATCH_MULTI", str(request.param))
    return request.param


def _dev(torch, *arrs):
    ts = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda() for a in arrs]
    torch.cuda.synchronize()
    return ts


def _single(fa, job, los, Nmu, cat):
    p = fa.PowerPlan(job["Ng"], job["L"], interlace=job.get("interlace", True),
                     dK=job.get("dK", 2 * np.pi / job["L"]), los=los, Nmu=Nmu,
                     Np_max=job["Np_max"], paint_tile=job.get("paint_tile", 0))
    out = p.run(*cat)
    p.close()
    return out


def _same(a, b, Pshot, tol=1e-12):
    assert np.array_equal(a["N"], b["N"])
    ok = a["N"] > 0
    assert np.allclose(a["K"][ok], b["K"][ok], rtol=1e-14, atol=0)
    scale = np.abs(b["Pl"][ok, 0]) + Pshot
    for l in range(4):
        err = np.abs(a["Pl"][ok, l] - b["Pl"][ok, l]) / scale
        assert err.max() <= tol, (l, err.max())
    if "Nmu" in b:
        assert np.array_equal(a["Nmu"], b["Nmu"])
        wm = b["Nmu"] > 0
        errw = np.abs(a["Pmu"][wm] - b["Pmu"][wm]) / (np.abs(b["Pmu"][wm]) + Pshot)
        assert errw.max() <= tol


def _catalogs(torch, specs):
    """specs: list of (Np, L, seed) -> list of (tensors, Np, V)."""
    from fftgal_amd import mockgen
    out = []
    for Np, L, seed in specs:
        (x, y, z), _ = mockgen.lognormal(Np, L, seed=seed, Ngs=32)
        out.append((_dev(torch, x, y, z), Np, L ** 3, (x, y, z)))
    return out


def _batch(fa, monkeypatch, sets, jobs, **kw):
    """A batch with `sets` int64 grid sets (2: the paint clears the other
    set; 1: the row FFT clears the grids)."""
    if sets == 1:
        monkeypatch.setenv("FFTGAL_GRID_SETS", "1")
    else:
        monkeypatch.delenv("FFTGAL_GRID_SETS", raising=False)
    b = fa.BatchPlan(jobs, **kw)
    monkeypatch.delenv("FFTGAL_GRID_SETS", raising=False)
    assert b.grid_sets() == sets
    return b


@pytest.mark.parametrize("sets", [2, 1])
def test_batch_matches_own_plans(torch_cuda, lib, monkeypatch, sets):
    """Uneven batch: grid sides 32/48/64/40 (specialised and generic FFT
    kernels), folds 1, 2 and 4 of one catalog, tiles 16 and 8, one catalog
    without interlacing, in one batch; each equals its own plan."""
    import fftgal_amd as fa
    torch = torch_cuda
    cats = _catalogs(torch, [(60000, 500.0, 1), (25000, 300.0, 2), (90000, 800.0, 3)])
    los, Nmu = (0.0, 0.0, 1.0), 5
    # (catalog, job)
    plan = [
        (0, {"Ng": 48, "L": 500.0}),
        (0, {"Ng": 48, "L": 250.0}),                     # fold 2
        (0, {"Ng": 32, "L": 125.0}),                     # fold 4
        (1, {"Ng": 32, "L": 300.0, "interlace": False}),
        (1, {"Ng": 64, "L": 300.0, "paint_tile": 8}),
        (2, {"Ng": 40, "L": 800.0}),                     # generic FFT kernels
        (2, {"Ng": 64, "L": 400.0, "dK": 2 * 2 * np.pi / 400.0}),
    ]
    jobs = []
    for c, j in plan:
        j = dict(j)
        j["Np_max"] = cats[c][1]
        jobs.append(j)
    b = _batch(fa, monkeypatch, sets, jobs, los=los, Nmu=Nmu)
    args = [(cats[c][0][0].data_ptr(), cats[c][0][1].data_ptr(), cats[c][0][2].data_ptr(),
             cats[c][1], cats[c][2]) for c, _ in plan]
    res = b.run(args)
    for q, (c, j) in enumerate(plan):
        ref = _single(fa, jobs[q], los, Nmu, args[q])
        print(q, j, "tile", b.tile(q))
        _same(res[q], ref, cats[c][2] / cats[c][1])
    b.close()


def test_batch_folded_vs_oracle(torch_cuda, lib, orc):
    """A catalog folded twice (grid box L/2, positions in box L) inside a
    batch, against the FP64 oracle: counts exact, P_l within 1e-5."""
    import fftgal_amd as fa
    torch = torch_cuda
    (t, Np, V, (x, y, z)), = _catalogs(torch, [(80000, 600.0, 7)])
    Lg, Ng, Nmu = 300.0, 48, 4
    los = (0.0, 0.0, 1.0)
    dK = 2 * np.pi / Lg
    b = fa.BatchPlan([{"Ng": Ng, "L": Lg, "Np_max": Np},
                      {"Ng": 32, "L": 600.0, "Np_max": Np}], los=los, Nmu=Nmu)
    cat = (t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), Np, V)
    out = b.run([cat, cat])[0]
    b.close()
    ref = orc.power_pipeline(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64),
                             V, Ng, Lg, interlace=True, los=los, dK=dK, Nmu=Nmu)
    assert np.array_equal(out["N"], ref["N"])
    assert np.array_equal(out["Nmu"], ref["Nmu"])
    ok = ref["N"] > 0
    sel = ok & (np.arange(len(ok)) < Ng // 2)
    scale = np.abs(ref["Pl"][sel, 0]) + V / Np
    for l in range(3):
        err = np.abs(out["Pl"][sel, l] - ref["Pl"][sel, l]) / scale
        assert err.max() <= 1e-5, (l, err.max())


@pytest.mark.parametrize("sets", [2, 1])
def test_batch_replays_and_new_catalogs(torch_cuda, lib, monkeypatch, sets):
    """The captured graphs replay; a new set of catalogs (other arrays,
    other Np and V) updates the tables and replays the same graphs.  Graph
    and eager runs and profiles interleave (with two grid sets each run
    paints one set and clears the other).  Each run must equal the plans of
    its own catalogs, and a rerun of the same catalogs must give the same
    bins bit for bit (integer grids, exact fixed-point bin sums)."""
    import fftgal_amd as fa
    torch = torch_cuda
    cats = _catalogs(torch, [(40000, 400.0, 11), (30000, 400.0, 12), (50000, 700.0, 13)])
    los, Nmu = (0.0, 1.0, 1.0), 3
    jobs = [{"Ng": 32, "L": 400.0, "Np_max": 50000}, {"Ng": 48, "L": 200.0, "Np_max": 50000}]
    b = _batch(fa, monkeypatch, sets, jobs, los=los, Nmu=Nmu)

    def args_of(c):
        t, Np, V, _ = cats[c]
        return (t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), Np, V)

    seen = {}
    for mode, sel in (("graph", (0, 1)), ("graph", (0, 1)), ("graph", (2, 1)),
                      ("profile", (0, 2)), ("graph", (1, 0)), ("eager", (2, 0)),
                      ("graph", (0, 1)), ("eager", (1, 2)), ("graph", (2, 1))):
        args = [args_of(c) for c in sel]
        if mode == "profile":
            b.profile(args)
            continue
        b.use_graph(mode == "graph")
        res = b.run(args)
        torch.cuda.synchronize()
        for q, c in enumerate(sel):
            ref = _single(fa, jobs[q], los, Nmu, args[q])
            _same(res[q], ref, cats[c][2] / cats[c][1])
            if sel in seen:
                for key in ("N", "K", "Pl", "Nmu", "Kmu", "Pmu"):
                    assert np.array_equal(res[q][key], seen[sel][q][key], equal_nan=True), key
        seen.setdefault(sel, res)
    b.close()


def test_batch_rejects_bad_input(torch_cuda, lib):
    import fftgal_amd as fa
    with pytest.raises(RuntimeError):
        fa.BatchPlan([{"Ng": 34, "L": 100.0, "Np_max": 10}])      # 17 is not a factor
    with pytest.raises(RuntimeError):
        fa.BatchPlan([{"Ng": 512, "L": 100.0, "Np_max": 10, "paint_tile": 8}])
    b = fa.BatchPlan([{"Ng": 16, "L": 100.0, "Np_max": 10}])
    with pytest.raises(RuntimeError):
        b.run([(0, 0, 0, 11, 1e6)])                                 # Np > Np_max
    b.close()


@pytest.mark.parametrize("serial", [0, 1], ids=["two_streams", "one_stream"])
@pytest.mark.parametrize("sets", [2, 1])
def test_batch_rounds_match_one_round(torch_cuda, lib, monkeypatch, sets, serial):
    """A device pool bound (FFTGAL_BATCH_BYTES) small enough to cut the
    catalogs into several rounds, which share the pool's grid sets, spectra
    and sort words: the bins equal those of the batch in one round bit for
    bit (the same integer grids and fixed-point sums), over graph replays
    and eager runs, with the grid sets alternating across rounds and
    enqueues.  Each round's sort runs on a second stream beside the previous
    round's FFT passes and bin, or (FFTGAL_BATCH_SERIAL=1) on the one
    stream."""
    if serial:
        monkeypatch.setenv("FFTGAL_BATCH_SERIAL", "1")
    else:
        monkeypatch.delenv("FFTGAL_BATCH_SERIAL", raising=False)
    import fftgal_amd as fa
    torch = torch_cuda
    cats = _catalogs(torch, [(60000, 500.0, 1), (25000, 300.0, 2), (90000, 800.0, 3)])
    los, Nmu = (0.0, 0.0, 1.0), 5
    plan = [(0, {"Ng": 48, "L": 500.0}), (0, {"Ng": 48, "L": 250.0}),
            (1, {"Ng": 32, "L": 300.0, "interlace": False}), (1, {"Ng": 64, "L": 300.0}),
            (2, {"Ng": 40, "L": 800.0}), (2, {"Ng": 64, "L": 400.0})]
    jobs = []
    for c, j in plan:
        j = dict(j)
        j["Np_max"] = cats[c][1]
        jobs.append(j)
    args = [(cats[c][0][0].data_ptr(), cats[c][0][1].data_ptr(), cats[c][0][2].data_ptr(),
             cats[c][1], cats[c][2]) for c, _ in plan]
    one = _batch(fa, monkeypatch, sets, jobs, los=los, Nmu=Nmu)
    assert one.rounds() == 1
    ref = one.run(args)
    one.close()
    bound = (16 if sets == 2 else 9) << 20
    monkeypatch.setenv("FFTGAL_BATCH_BYTES", str(bound))
    many = _batch(fa, monkeypatch, sets, jobs, los=los, Nmu=Nmu)
    monkeypatch.delenv("FFTGAL_BATCH_BYTES", raising=False)
    print("rounds", many.rounds(), "pool MiB", many.pool_bytes() / 2 ** 20)
    assert many.rounds() >= 3
    assert many.pool_bytes() <= bound
    for mode in ("graph", "graph", "eager", "graph"):
        many.use_graph(mode == "graph")
        res = many.run(args)
        for q in range(len(plan)):
            for key in ("N", "K", "Pl", "Nmu", "Kmu", "Pmu"):
                assert np.array_equal(res[q][key], ref[q][key], equal_nan=True), (mode, q, key)
    many.close()

"""Benchmark: Mparticles/s to P_l(k) on one MI355X.

Workload (default): lognormal mock of 134,217,728 float32 particles in a
periodic L = 2000 Mpc/h box (power-law P(k) = 2e4 (k/0.1)^-1.5 on a 256^3
seed grid, Zel'dovich RSD along z with f = 0.76), Ng = 512, PCS paint with
interlacing, multipoles l = 0,2,4 (6 computed too) and 10 |mu| wedges,
dK = kF (NK = 444 bins, 256 below k_Nyquist).

A step is one catalog through the mixed-precision pipeline (fixed-point
paint of both grids, two FP32 r2c FFTs, fused deconvolve/interlace/bin,
bin sums to host), positions already in HBM, replayed as a hipGraph.
Also measured in the same run:
  * per-stage times (HIP events on the pipeline's stream) and the HBM
    roofline of the dominant kernel;
  * the plain FP64 GPU form (drop-in API, device-resident, FP64 atomics
    and FP64 FFT) on the same catalog, and the P_l error of the fast path
    against it;
  * the FP64 CPU oracle (restatement of the reference) on the whole
    catalog with its threads (the CPU baseline), the error of the fast path
    against it, and a one-thread estimate composed of measured parts.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Mparticles/s to P_l(k) at Ng=512, batched mixed vs plain FP64; "
          "HBM GB/s vs peak")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=(3, 4, 5),
                    help="3: 2^27 particles, Ng=512, L=2000, 10 wedges (default); "
                         "4: batch of 64 lognormal mocks (L 250-2000, Ng 128-384, "
                         "5e5-2e7 particles) each at folds 1, 2, 4 (192 catalogs); "
                         "5: 2^30 particles, Ng=2048, L=4000, 20 wedges (large single "
                         "grid; the CPU baseline is skipped there)")
    ap.add_argument("--mocks", type=int, default=64, help="config 4: number of mocks")
    ap.add_argument("--workers", type=int, default=16,
                    help="config 4: processes generating the mocks")
    ap.add_argument("--fold", type=int, default=1,
                    help="grid box L/fold with the catalog in box L (src/app/P.c:35-36)")
    ap.add_argument("--np", type=int, default=None)
    ap.add_argument("--ng", type=int, default=None)
    ap.add_argument("--L", type=float, default=None)
    ap.add_argument("--nmu", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 21)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default: tune each grid side on the workload "
                         "first, timed separately)")
    ap.add_argument("--extras", action="store_true",
                    help="time the off-path drop-in operations (x2k, k2x, randoms paint, "
                         "super-sample filter, window deconvolution, QPM loaders) instead; "
                         "one JSON line per operation")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed graph loop (for rocprofv3 runs)")
    a = ap.parse_args()
    preset = {3: (1 << 27, 512, 2000.0, 10), 4: (0, 0, 0.0, 10),
              5: (1 << 30, 2048, 4000.0, 20)}[a.config]
    a.np = a.np if a.np is not None else preset[0]
    a.ng = a.ng if a.ng is not None else preset[1]
    a.L = a.L if a.L is not None else preset[2]
    a.nmu = a.nmu if a.nmu is not None else preset[3]
    if a.config == 5:
        a.no_cpu = True
    return a


def stage_bytes(name, Np, Ng, sets=1):
    """Bytes one stage of one catalog moves as designed (per-stage GB/s).
    The int64 grids are cleared for the next paint by the row FFT (one grid
    set) or by the paint itself, which clears the other set (two sets)."""
    n3 = Ng ** 3
    spec = 4 * n3                                # FP32 packed half spectrum: Ng^3/2 complex
    clear = 2 * 8 * n3                           # both int64 grids
    table = {
        "sort1_count": 12 * Np,                 # read float32 xyz
        # read xyz, write the 8-byte main word and the 2-byte side word
        "sort1_scatter": 12 * Np + 10 * Np,
        "sort2_count": 2 * Np,                  # read the side words
        # read main + side words, write main word + 1-byte extension
        "sort2_scatter": 10 * Np + 9 * Np,
        # read the packed positions, RMW both int64 grids once (and clear
        # the other set)
        "paint_tiles": 9 * Np + 2 * 16 * n3 + (clear if sets == 2 else 0),
        # both grids: read the int64 grid, write the spectrum (and clear the
        # grid for the next paint)
        "fft_rows": 2 * (8 * n3 + spec) + (clear if sets == 1 else 0),
        "fft_cols_j": 2 * 2 * spec, "fft_cols_i": 2 * 2 * spec,   # read + write FP32
        "fused_bin": 2 * spec,                  # read both spectra
    }
    return table.get(name, 0)


def paint_compulsory(Np, Ng, G=2, s=4, bp=12):
    """Compulsory bytes of the mass assignment: read the positions once and
    write each FP32 grid once, b_p Np + G s Ng^3 (the first two terms of the
    pipeline's B_alg).  Sort words and int64 scratch are overhead."""
    return bp * Np + G * s * Ng ** 3


def b_alg(Np, Ng, G=2, s=4, bp=12):
    """Compulsory bytes of the whole pipeline: b_p Np + G s (Ng^3 + 3 Nc)."""
    return bp * Np + G * s * (Ng ** 3 + 3 * Ng * Ng * (Ng + 2))


def copy_bandwidth(torch, nbytes=1 << 30, reps=5):
    """Measured device copy rate (bytes read + written per second, GB/s) of
    a plain 1 GiB copy, the achievable HBM figure beside the 8 TB/s peak."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fp:
        return json.load(fp)


def cpu_baseline(x, y, z, V, Ng, L, nmu, nsample):
    """The FP64 oracle (the reference's algorithm restated in C) on the CPU:
      * measured: the whole catalog through the reference's call order
        (two paints, two FFTs, average, deconv, P, Pl, Pmu) with the
        oracle's threads (its results are bit-identical to one thread);
      * a one-thread estimate like the reference's own single-threaded run,
        composed of measured parts: every grid-sized stage on the full grid,
        and the paint's particle loop on a sample of nsample particles,
        scaled by Np / nsample (the paint's grid sweeps, zeroing and
        normalising, are measured separately and not scaled).
    Returns the baseline dict and the oracle's bins (for error_vs_oracle)."""
    import numpy as np
    import oracle
    Np = len(x)
    dK = 2 * np.pi / L
    threads = oracle.default_threads()
    xd, yd, zd = (a.astype(np.float64) for a in (x, y, z))
    oracle.set_threads(threads)
    try:
        t = {}
        t0 = time.perf_counter()
        ref = oracle.power_pipeline(xd, yd, zd, V, Ng, L, interlace=True, los=(0, 0, 1), dK=dK,
                                    Nmu=nmu, timings=t)
        t_full = time.perf_counter() - t0
    finally:
        oracle.set_threads(1)
    print(f"# oracle, {threads} threads, full catalog: {t_full:.1f}s", file=sys.stderr, flush=True)
    # one thread: grid-sized stages on the full grid, particle loop on a sample
    n = min(nsample, Np)
    st = {}
    small = oracle.Catalog(xd[:1], yd[:1], zd[:1], V)
    samp = oracle.Catalog(xd[:n], yd[:n], zd[:n], V)
    g = oracle.Grid(Ng, L)
    t0 = time.perf_counter()
    g.p2g(small, (0.5, 0.5, 0.5))
    st["p2g_sweeps"] = time.perf_counter() - t0          # zero + normalise
    t0 = time.perf_counter()
    g.p2g(samp, (0.5, 0.5, 0.5))
    st["p2g_sample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.x2k(1)
    st["x2k"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.deconv()
    g.P(samp)
    g.Pl_bins((0, 0, 1), dK)
    g.Pmu_bins((0, 0, 1), dK, nmu)
    st["deconv_P_Pl_Pmu"] = time.perf_counter() - t0
    g.free()
    small.free()
    samp.free()
    loop = max(st["p2g_sample"] - st["p2g_sweeps"], 0.0) * (Np / n)
    # two paints and two FFTs (interlacing), the average sweep ~ one grid pass
    t_serial = 2 * (loop + st["p2g_sweeps"]) + 2 * st["x2k"] + st["p2g_sweeps"] + \
        st["deconv_P_Pl_Pmu"]
    base = {
        "value": round(Np / t_full / 1e6, 4),
        "unit": "Mparticles/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the whole catalog ({Np} particles, Ng={Ng}) through the FP64 oracle in "
                   f"the reference's call order (2 paints, 2 FFTs, average, deconv, P, Pl, "
                   f"Pmu), {threads} threads, measured: {t_full:.1f}s"),
        "stage_s": {k: round(v, 3) for k, v in t.items()},
        "one_thread_estimate": {
            "Mparticles_per_s": round(Np / t_serial / 1e6, 4), "s": round(t_serial, 1),
            "method": (f"one thread: grid stages measured on the full grid; particle loop "
                       f"measured on {n} particles (minus the grid sweeps) and scaled by "
                       f"{Np / n:.0f}"),
            "stage_s": {k: round(v, 3) for k, v in st.items()}},
        "cpu": _cpu_model(),
    }
    return base, ref


def parity(out, ref, Ng, Pshot, eps=0.05, eps_peak=1e-3):
    """Error of the fast path against the oracle: the largest |dP/P| over
    bins with |P| >= eps (|P_0| + P_shot) + eps_peak P_peak (k < k_Nyquist),
    and the largest |dP| / (|P_0| + P_shot) over all of them."""
    import numpy as np
    nb = len(ref["N"])
    sel = (ref["N"] > 0) & (np.arange(nb) < Ng // 2)
    P0 = np.abs(ref["Pl"][sel, 0])
    keep = eps * (P0 + Pshot) + eps_peak * P0.max()
    res = {"counts_equal": bool(np.array_equal(out["N"], ref["N"]) and
                                ("Nmu" not in ref or np.array_equal(out["Nmu"], ref["Nmu"])))}
    for l in range(3):
        r, o = ref["Pl"][sel, l], out["Pl"][sel, l]
        d = np.abs(o - r)
        big = np.abs(r) >= keep
        res[f"P{2 * l}"] = {"max_rel_kept": float((d[big] / np.abs(r[big])).max()) if big.any()
                            else None, "kept_bins": int(big.sum()), "bins": int(sel.sum()),
                            "max_abs_over_P0_Pshot": float((d / (P0 + Pshot)).max())}
    if "Pmu" in ref:
        w = (ref["Nmu"] > 0) & (np.arange(nb) < Ng // 2)[:, None]
        p0 = (np.abs(ref["Pl"][:, 0])[:, None] * np.ones_like(ref["Pmu"]))[w]
        r, o = ref["Pmu"][w], out["Pmu"][w]
        d = np.abs(o - r)
        big = np.abs(r) >= eps * (p0 + Pshot) + eps_peak * P0.max()
        res["wedges"] = {"max_rel_kept": float((d[big] / np.abs(r[big])).max()),
                         "kept_bins": int(big.sum()), "bins": int(w.sum()),
                         "max_abs_over_P0_Pshot": float((d / (p0 + Pshot)).max())}
    res["rule"] = (f"k < k_Nyquist; kept: |P| >= {eps} (|P0| + Pshot) + {eps_peak} P_peak; "
                   "tests require max_rel_kept <= 1e-5")
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _batch_mock(spec):
    from fftgal_amd import mockgen
    s, L, Ng, Np = spec
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=s, Ngs=256)
    return x, y, z


def main_batch(args):
    """Batched uneven workload: every mock at folds 1, 2 and 4, all
    catalogs through one fftgal_batch_t (one launch per stage, one graph),
    against the plain FP64 form run catalog by catalog."""
    import multiprocessing as mp
    import numpy as np
    from fftgal_amd.mockgen import batch_specs
    specs = batch_specs(args.mocks)
    t0 = time.perf_counter()
    # the mocks are drawn before anything touches the GPU (forked workers)
    with mp.get_context("fork").Pool(max(1, args.workers)) as pool:
        mocks = pool.map(_batch_mock, specs, chunksize=1)
    t_mock = time.perf_counter() - t0
    print(f"# {len(mocks)} mocks, {sum(s[3] for s in specs)} particles in {t_mock:.1f}s",
          file=sys.stderr, flush=True)
    import torch
    import fftgal_amd as fa
    fa.load()
    nmu = args.nmu if args.nmu is not None else 10
    folds = (1, 2, 4)
    dev = []
    for (x, y, z) in mocks:
        dev.append([torch.from_numpy(a).cuda() for a in (x, y, z)])
    torch.cuda.synchronize()
    jobs, cats = [], []
    for (s, L, Ng, Np), d in zip(specs, dev):
        for fold in folds:
            Lg = L / fold
            jobs.append({"Ng": Ng, "L": Lg, "Np_max": Np, "dK": 2 * np.pi / Lg})
            cats.append((d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), Np, L ** 3))
    Ntot = sum(c[3] for c in cats)
    wisdom, tuned, t_tune = None, {}, 0.0
    if not args.no_tune:
        import tempfile
        wisdom = os.path.join(tempfile.mkdtemp(prefix="fftgal_tune_"), "batch.wsdm")
        t0 = time.perf_counter()
        for n in sorted({sp[2] for sp in specs}):
            q = next(i for i, sp in enumerate(specs) if sp[2] == n)
            s_, L_, _, Np_ = specs[q]
            d = dev[q]
            tuned[n] = fa.autotune(n, L_, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), Np_,
                                   L_ ** 3, wisdom, Nmu=nmu, reps=2)
        t_tune = time.perf_counter() - t0
        print(f"# autotune: {tuned} in {t_tune:.1f}s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t0 = time.perf_counter()
    b = fa.BatchPlan(jobs, los=(0, 0, 1), Nmu=nmu, wisdom=wisdom)
    t_plan = time.perf_counter() - t0
    batch_bytes = free0 - torch.cuda.mem_get_info()[0]   # device memory the batch took
    rounds, pool = b.rounds(), b.pool_bytes()
    print(f"# batch plan: {len(jobs)} catalogs in {t_plan:.1f}s, {rounds} rounds, pool "
          f"{pool / 2 ** 30:.1f} GiB, device memory taken {batch_bytes / 1e9:.1f} GB",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        b.enqueue(cats)
    b.sync()
    stream = torch.cuda.ExternalStream(b.stream())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        b.enqueue(cats)
    e1.record(stream)
    b.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    value = Ntot / (ms_per_step * 1e-3) / 1e6
    print(f"# batched: {ms_per_step:.2f} ms per batch ({value:.0f} Mparticles/s)",
          file=sys.stderr, flush=True)
    fast = [b.results(j) for j in range(len(jobs))]
    prof = b.profile(cats)
    sets = b.grid_sets()
    b.close()
    torch.cuda.empty_cache()
    if args.profile_only:   # for kernel traces: no plain FP64 comparison
        print(json.dumps({"ms_per_step": ms_per_step, "value": value, "grid_sets": sets,
                          "rounds": rounds, "pool_bytes": pool, "device_bytes": batch_bytes,
                          "stages_ms": {k: round(v, 3) for k, v in prof.items()}}))
        return

    # plain FP64 form, catalog by catalog (device resident)
    t_plain = 0.0
    worst = {"P0": 0.0, "P2": 0.0, "P4": 0.0, "wedges": 0.0}
    where = {}
    counts_equal = True
    for q, ((s, L, Ng, Np), fold) in enumerate(
            [(sp, f) for sp in specs for f in folds]):
        d = dev[q // len(folds)]
        dd = [t.double() for t in d]
        torch.cuda.synchronize()
        Lg = L / fold
        g = fa.Grid(Ng, Lg, resident=True)
        lib = g.lib
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lib.fft_p2k_device(g.ptr, dd[0].data_ptr(), dd[1].data_ptr(), dd[2].data_ptr(), Np,
                           L ** 3)
        lib.P_params(g.ptr, Np, L ** 3, 0)
        K, N, Pl = g.Pl_bins((0, 0, 1), 2 * np.pi / Lg)
        Km, Nm, Pm = g.Pmu_bins((0, 0, 1), 2 * np.pi / Lg, nmu)
        t_plain += time.perf_counter() - t1
        g.free()
        del dd
        o = fast[q]
        counts_equal &= bool(np.array_equal(o["N"], N) and np.array_equal(o["Nmu"], Nm))
        sel = (N > 0) & (np.arange(len(N)) < Ng // 2)
        Pshot = L ** 3 / Np
        for l, key in enumerate(("P0", "P2", "P4")):
            # relative to |P_l|, or to |P_0| + Pshot where P_l is near zero
            scale = np.maximum(np.abs(Pl[sel, l]), np.abs(Pl[sel, 0]) + Pshot)
            rel = np.abs(o["Pl"][sel, l] - Pl[sel, l]) / scale
            e = float(rel.max())
            if e > worst[key]:
                b = int(np.flatnonzero(sel)[int(rel.argmax())])
                where[key] = {"catalog": q, "Ng": Ng, "L_grid": round(Lg, 2), "Np": Np,
                              "bin": b, "N": int(N[b]), "fast": float(o["Pl"][b, l]),
                              "plain": float(Pl[b, l]), "P0": float(Pl[b, 0]),
                              "Pshot": Pshot}
            worst[key] = max(worst[key], e)
        wm = (Nm[:Ng // 2] > 0)
        ew = np.abs(o["Pmu"][:Ng // 2][wm] - Pm[:Ng // 2][wm]) / (np.abs(Pm[:Ng // 2][wm]) + Pshot)
        worst["wedges"] = max(worst["wedges"], float(ew.max()))
    torch.cuda.empty_cache()
    plain_value = Ntot / t_plain / 1e6
    print(f"# plain FP64: {t_plain * 1e3:.0f} ms for all catalogs", file=sys.stderr, flush=True)
    cells = sum(j["Ng"] ** 3 for j in jobs)
    line = {
        "metric": "Mparticles/s to P_l(k), batched uneven workload (192 catalogs), "
                  "mixed vs plain FP64",
        "value": round(value, 1), "unit": "Mparticles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f32/f64",
        "data": ("64 synthetic lognormal mocks (numpy, seeds 0..63, 256^3 seed grids, "
                 "Zel'dovich RSD along z), each at folds 1, 2, 4"),
        "config": {"workload": f"{len(jobs)} catalogs (64 mocks x folds 1,2,4), "
                               f"sum Np = {Ntot}, sum Ng^3 = {cells}, interlaced, "
                               f"l=0,2,4(,6) and {nmu} |mu|-wedges, dK=kF of each grid box",
                   "grids": {str(n): sum(1 for j in jobs if j["Ng"] == n)
                             for n in (128, 192, 256, 384)},
                   "grid_sets": sets, "rounds": rounds, "pool_bytes": pool,
                   "device_bytes": batch_bytes, "parallelism": "1 GPU"},
        "stages_ms": {k: round(v, 3) for k, v in prof.items()},
        "plain_fp64": {"value": round(plain_value, 2), "unit": "Mparticles/s",
                       "ms_total": round(t_plain * 1e3, 1),
                       "form": "drop-in API per catalog, device resident, FP64"},
        "speedup_vs_plain_fp64": round(value / plain_value, 2),
        "error_vs_plain_fp64": {"max_rel_dP_below_kNyq": worst,
                                "relative_to": "max(|P_l|, |P_0| + Pshot) of the bin",
                                "counts_equal": counts_equal, "worst": where},
        "timing": {"mock_s": round(t_mock, 1), "autotune_s": round(t_tune, 2),
                   "plan_s": round(t_plan, 2), "wall_s": round(wall, 3)},
        "tuning": {str(k): v for k, v in tuned.items()},
    }
    print(json.dumps(line))


def main_extras(args):
    """Times of the drop-in API's operations beside the power-spectrum path,
    on an Ng = 512, L = 2000 grid kept on the device (fft_set_device_resident):
    fft_x2k / fft_k2x (FP64 R2C / C2R with H^3, 1/L^3 and the offset phase),
    fft_p2g of a catalog with randoms (FP64 atomics), fft_dl_filter (filter,
    C2R, 4^3 samples) and fft_subvolume_deconv (C2R, window, R2C), and the
    QPM loaders on a synthetic file of 2^21 lines.  Wall time per call with
    the device synchronised; a grid operation's compulsory bytes are one
    read and one write of the padded FP64 grid per pass over it."""
    import ctypes as C
    import tempfile
    import numpy as np
    import torch
    import fftgal_amd as fa
    lib = fa.load()
    Ng, L = 512, 2000.0
    Nc = Ng * Ng * 2 * (Ng // 2 + 1)
    rw = 2 * 8 * Nc
    reps = max(2, min(args.steps, 5))
    dp = C.POINTER(C.c_double)

    def emit(op, secs, alg_bytes, **kw):
        s = sorted(secs)[len(secs) // 2]
        line = {"op": op, "ms": round(s * 1e3, 3), "reps": len(secs), "alg_bytes": alg_bytes,
                "GBps": round(alg_bytes / s / 1e9, 1) if alg_bytes else None,
                "frac_of_hbm_peak": round(alg_bytes / s / 1e9 / HBM_PEAK_GBS, 4)
                if alg_bytes else None}
        line.update(kw)
        print(json.dumps(line), flush=True)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    rng = np.random.default_rng(1)
    g = fa.Grid(Ng, L, resident=True)
    n, nr = 1 << 24, 1 << 25
    cat = fa.Catalog(*(rng.random((3, n)) * L), L ** 3)
    ran = fa.Catalog(*(rng.random((3, nr)) * L), L ** 3)
    cat.ptr.contents.rand = ran.ptr
    ran.ptr = None                       # freed with its parent
    off = np.zeros(3)
    secs = [timed(lambda: lib.fft_p2g(g.ptr, cat.ptr, off.ctypes.data_as(dp)))
            for _ in range(reps + 1)][1:]
    emit("fft_p2g with randoms", secs, 24 * (n + nr) + rw // 2, Np=n, Nrand=nr,
         bound="FP64 global atomics (64 per particle)")
    t_x, t_k = [], []
    for r in range(reps + 1):
        a = timed(lambda: g.x2k(1))
        b = timed(lambda: g.k2x(1))
        if r:
            t_x.append(a)
            t_k.append(b)
    emit("fft_x2k (phase on)", t_x, 3 * rw, passes="rows, columns j, columns i")
    emit("fft_k2x (phase off)", t_k, 3 * rw, passes="columns i, columns j, rows")
    g.x2k(0)
    los = np.array([0.0, 0.0, 1.0])
    delta = np.zeros(64)
    secs = [timed(lambda: lib.fft_dl_filter(g.ptr, 0, 4, los.ctypes.data_as(dp), 2, 1.0,
                                            delta.ctypes.data_as(dp)))
            for _ in range(reps + 1)][1:]
    emit("fft_dl_filter (sub-box, Nsub 4, l 2)", secs, 4 * rw)
    secs = [timed(lambda: lib.fft_subvolume_deconv(g.ptr, 0, 4)) for _ in range(reps + 1)][1:]
    emit("fft_subvolume_deconv (sub-box, Nsub 4)", secs, 7 * rw)
    g.free()
    cat.free()
    # QPM loaders: x y z vx vy vz halo_mass sat_flag halo_id, 2^21 lines
    nq = 1 << 21
    fd, path = tempfile.mkstemp(suffix=".qpm")
    os.close(fd)
    try:
        cols = [rng.random(nq) * 1000.0 for _ in range(3)] + \
            [rng.standard_normal(nq) * 300.0 for _ in range(3)]
        with open(path, "w") as fp:
            for i0 in range(0, nq, 1 << 18):
                sl = slice(i0, i0 + (1 << 18))
                block = np.column_stack([c[sl] for c in cols] + [np.full(len(cols[0][sl]), 1e13),
                                                                 np.zeros(len(cols[0][sl])),
                                                                 np.arange(len(cols[0][sl]))])
                np.savetxt(fp, block, fmt="%.6f %.6f %.6f %.4f %.4f %.4f %.6e %d %d")
        size = os.path.getsize(path)
        secs = []
        for _ in range(reps):
            t0 = time.perf_counter()
            gp = lib.gal_loadqpm_cubic(path.encode(), 1000.0)
            secs.append(time.perf_counter() - t0)
            lib.gal_free(gp)
        emit("gal_loadqpm_cubic (host)", secs, 0, lines=nq, file_bytes=size,
             MBps=round(size / sorted(secs)[len(secs) // 2] / 1e6, 1))
        xs = [torch.empty(nq, dtype=torch.float32, device="cuda:0") for _ in range(3)]
        got = C.c_long()
        secs = []
        for _ in range(reps):
            secs.append(timed(lambda: lib.fftgal_qpm_load_device(
                path.encode(), nq, xs[0].data_ptr(), xs[1].data_ptr(), xs[2].data_ptr(),
                C.byref(got))))
        assert got.value == nq, got.value
        emit("fftgal_qpm_load_device", secs, 0, lines=nq, file_bytes=size,
             MBps=round(size / sorted(secs)[len(secs) // 2] / 1e6, 1))
    finally:
        os.unlink(path)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (independent "
              "catalogs would shard across GPUs, one process each, with no communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.extras:
        return main_extras(args)
    if args.config == 4:
        return main_batch(args)
    import numpy as np
    import torch
    import fftgal_amd as fa
    from fftgal_amd import mockgen

    fa.load()
    Np, Ng, L, nmu = args.np, args.ng, args.L, args.nmu
    V = L ** 3                 # catalog volume
    Lg = L / args.fold         # grid box (folding)
    dK = 2 * np.pi / Lg
    t0 = time.perf_counter()
    (x, y, z), sigma2 = mockgen.lognormal(
        Np, L, seed=0, Ngs=256,
        log=lambda m: print(f"# {m} ({time.perf_counter() - t0:.0f}s)", file=sys.stderr,
                            flush=True))
    t_mock = time.perf_counter() - t0
    print(f"# mock: {Np} particles in {t_mock:.1f}s (sigma_G^2 = {sigma2:.2f})",
          file=sys.stderr, flush=True)
    dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
    torch.cuda.synchronize()
    ptr = [d.data_ptr() for d in dev]

    t0 = time.perf_counter()
    wisdom, tuned, t_tune = None, None, 0.0
    if not args.no_tune:
        import tempfile
        wisdom = os.path.join(tempfile.mkdtemp(prefix="fftgal_tune_"), "bench.wsdm")
        tuned = fa.autotune(Ng, Lg, *ptr, Np, V, wisdom, dK=dK, Nmu=nmu, reps=2)
        t_tune = time.perf_counter() - t0
        print(f"# autotune: {tuned} in {t_tune:.1f}s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
    plan = fa.PowerPlan(Ng, Lg, interlace=True, dK=dK, los=(0, 0, 1), Nmu=nmu, Np_max=Np,
                        wisdom=wisdom)
    t_plan = time.perf_counter() - t0
    print(f"# plan: {t_plan:.2f}s", file=sys.stderr, flush=True)
    stream = torch.cuda.ExternalStream(plan.stream())

    for _ in range(args.warmup):
        plan.enqueue(*ptr, Np, V)
    plan.sync()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.enqueue(*ptr, Np, V)
    ev1.record(stream)
    plan.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    fast = plan.results()
    value = Np / (ms_per_step * 1e-3) / 1e6
    print(f"# fast: {ms_per_step:.3f} ms/catalog (events {ev_ms:.3f}) -> {value:.1f} Mparticles/s",
          file=sys.stderr, flush=True)
    if args.profile_only:
        # calibration for HBM counters: a plain device copy of known size
        nbytes = 1 << 30
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"ms_per_step": ms_per_step, "value": value,
                          "copy_calibration_bytes": {"read": nbytes, "write": nbytes}}))
        return

    # per-stage times, HIP events on the pipeline's stream
    sets = plan.grid_sets()
    nprof = max(3, min(args.steps, 5))
    acc = {}
    for _ in range(nprof):
        st = plan.profile(*ptr, Np, V)
        for k, v in st.items():
            acc[k] = acc.get(k, 0.0) + v / nprof
    print("# stage profile done", file=sys.stderr, flush=True)
    stages = {}
    for k, ms in acc.items():
        b = stage_bytes(k, Np, Ng, sets)
        stages[k] = {"ms": round(ms, 4), "alg_bytes": b,
                     "GBps": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b else None}
    dom = max((k for k in acc if stage_bytes(k, Np, Ng, sets)), key=lambda k: acc[k])
    dom_ms = acc[dom]
    # the dominant kernel is the paint: its compulsory share of the pipeline's
    # bytes (positions read once, both FP32 grids written once) over its time;
    # what its design moves beyond that (sorted words, int64 read-modify-write,
    # the clear of the other grid set) is reported as overhead
    comp = paint_compulsory(Np, Ng) if dom == "paint_tiles" else stage_bytes(dom, Np, Ng, sets)
    ach = comp / (dom_ms * 1e-3) / 1e9
    traffic = None
    tr = load_traffic()
    # the counter profile counts only for the workload it was taken on
    wl = (tr or {}).get("workload", {})
    if tr and dom in tr.get("stages", {}) and wl.get("Np") == Np and wl.get("Ng") == Ng:
        traffic = tr["stages"][dom].get("bytes")
    copy_gbs = copy_bandwidth(torch)
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "kernel_ms": round(dom_ms, 4), "compulsory_bytes": comp,
                "overhead_bytes": stage_bytes(dom, Np, Ng, sets) - comp + 12 * Np,
                "traffic_over_compulsory": round(traffic / comp, 3) if traffic else None,
                "copy_GBps": round(copy_gbs, 1),
                "frac_of_copy": round(ach / copy_gbs, 4),
                "note": ("achieved = compulsory bytes (12 Np + 2 x 4 Ng^3) / paint kernel time; "
                         "overhead_bytes = what the paint kernel moves beyond its own share "
                         "(packed positions, int64 grid read-modify-write, clear of the other "
                         "grid set); copy_GBps = a measured 1 GiB device copy, read + write")}
    # paint stencil updates: 64 cells per particle and grid (PCS), both grids
    upd = 64 * 2 * Np
    paint = {"stencil_updates": upd,
             "updates_per_s": round(upd / (acc["paint_tiles"] * 1e-3), -6),
             "kernel_ms": round(acc["paint_tiles"], 4),
             "note": "int32 LDS atomics; the footprints then flush with int64 global atomics"}
    B = b_alg(Np, Ng)
    pipe = {"B_alg_bytes": B, "GBps": round(B / (ms_per_step * 1e-3) / 1e9, 1),
            "frac_of_peak": round(B / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "stage_sum_ms": round(sum(acc.values()), 4)}
    plan_tuning = plan.tuning()
    plan.close()

    # plain FP64 GPU form on the same catalog
    plain = None
    err = None
    if not args.no_plain:
        dd = [d.double() for d in dev]
        torch.cuda.synchronize()
        g = fa.Grid(Ng, Lg, resident=True)
        lib = g.lib
        times = []
        for _ in range(args.plain_steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lib.fft_p2k_device(g.ptr, dd[0].data_ptr(), dd[1].data_ptr(), dd[2].data_ptr(), Np, V)
            lib.P_params(g.ptr, Np, V, 0)
            Kp, Npl, Pp = g.Pl_bins((0, 0, 1), dK)
            Km, Nm, Pm = g.Pmu_bins((0, 0, 1), dK, nmu)
            times.append(time.perf_counter() - t0)
        g.free()
        del dd
        tp = min(times)
        print(f"# plain FP64: {tp * 1e3:.1f} ms/catalog", file=sys.stderr, flush=True)
        plain = {"value": round(Np / tp / 1e6, 2), "unit": "Mparticles/s",
                 "ms_per_catalog": round(tp * 1e3, 2),
                 "form": "drop-in API, device resident: FP64 atomic paint x2, FP64 FFT x2, "
                         "phase, average, deconv, P, Pl, Pmu kernels"}
        err = parity(fast, {"N": Npl, "Pl": Pp, "Nmu": Nm, "Pmu": Pm}, Ng, V / Np)
    del dev
    torch.cuda.empty_cache()

    cpu, err_oracle = None, None
    if not args.no_cpu:
        cpu, ref = cpu_baseline(x, y, z, V, Ng, Lg, nmu, args.cpu_sample)
        err_oracle = parity(fast, ref, Ng, V / Np)
        err_oracle["strict_eps_0.02"] = parity(fast, ref, Ng, V / Np, eps=0.02)

    name = torch.cuda.get_device_name(0)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic lognormal mock (numpy, seed 0): P(k)=2e4 (k/0.1)^-1.5 on a 256^3 "
                 "seed grid, Poisson-sampled, Zel'dovich RSD along z f=0.76, random order"),
        "config": {"workload": (f"lognormal mock, {Np} float32 particles, L={L:g} Mpc/h, "
                                f"Ng={Ng}" + (f", fold {args.fold}" if args.fold > 1 else "") +
                                f", PCS paint + interlacing, l=0,2,4(,6) and "
                                f"{nmu} |mu|-wedges, dK=kF"),
                   "Np": Np, "Ng": Ng, "L": L, "fold": args.fold, "interlace": True,
                   "Nmu": nmu, "grid_sets": sets,
                   "precision": "int64 fixed-point paint, FP32 FFT (FP64 twiddles), "
                                "FP64 k-space math, exact fixed-point bin sums",
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "paint_stencil": paint,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 2) if plain else None,
        "error_vs_oracle": err_oracle,
        "error_vs_plain_fp64": err,
        "pipeline": pipe,
        "stages": stages,
        "event_ms_per_step": round(ev_ms, 4),
        "setup_s": {"mock": round(t_mock, 1), "plan": round(t_plan, 2),
                    "autotune": round(t_tune, 2)},
        "tuning": plan_tuning,
        "device": name,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

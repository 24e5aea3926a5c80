"""GPU tests of the autotuner (fftgal_power_autotune) and of every tunable
it picks from: whatever paint tile, FFT kernels and bin block the wisdom
file names, a plan must give the same mode counts and P_l within the fast
path's tolerance of the FP64 oracle (the bin block only changes the order of
the FP64 sums: 1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cat(torch, Np, L, seed):
    from fftgal_amd import mockgen
    (x, y, z), _ = mockgen.lognormal(Np, L, seed=seed, Ngs=32)
    t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z)]
    torch.cuda.synchronize()
    return (x, y, z), t


def _write_wisdom(path, Ng, interlace, tile, cols, rows, generic, bk):
    with open(path, "w") as fp:
        fp.write("# Ng interlace paint_tile fft_cols fft_rows fft_generic bin_bk\n")
        fp.write(f"{Ng} {interlace} {tile} {cols} {rows} {generic} {bk}\n")


def _rows(d, sel):
    return {k: v[sel] for k, v in d.items() if isinstance(v, np.ndarray) and len(v) == len(sel)}


def _close(a, b, Pshot, tol):
    assert np.array_equal(a["N"], b["N"])
    ok = b["N"] > 0
    scale = np.abs(b["Pl"][ok, 0]) + Pshot
    for l in range(3):
        assert (np.abs(a["Pl"][ok, l] - b["Pl"][ok, l]) / scale).max() <= tol
    if "Nmu" in b:
        assert np.array_equal(a["Nmu"], b["Nmu"])


def test_autotune_stores_and_plans_use_it(torch_cuda, lib, orc, tmp_path):
    import fftgal_amd as fa
    torch = torch_cuda
    Ng, Np, L, Nmu = 64, 150000, 600.0, 4
    (x, y, z), t = _cat(torch, Np, L, 21)
    V = L ** 3
    w = str(tmp_path / "tune.wsdm")
    ptr = [a.data_ptr() for a in t]
    best = fa.autotune(Ng, L, *ptr, Np, V, w, Nmu=Nmu, reps=1)
    print(best)
    lines = [ln.split() for ln in open(w) if not ln.startswith("#")]
    assert len(lines) == 1 and lines[0][:3] == ["64", "1", str(best["paint_tile"])]
    assert best["paint_tile"] in (16, 12, 8, 4) and best["bin_bk"] in (16, 32, 64, 128)
    p = fa.PowerPlan(Ng, L, dK=2 * np.pi / L, Nmu=Nmu, Np_max=Np, wisdom=w)
    tu = p.tuning()
    assert tu["paint_tile"] == best["paint_tile"] and tu["bin_bk"] == best["bin_bk"]
    out = p.run(*ptr, Np, V)
    p.close()
    ref = orc.power_pipeline(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64),
                             V, Ng, L, interlace=True, los=(0, 0, 1), dK=2 * np.pi / L, Nmu=Nmu)
    sel = ref["N"] > 0
    sel &= np.arange(len(sel)) < Ng // 2
    _close(_rows(out, sel), _rows(ref, sel), V / Np, 1e-5)
    # a second tuning of another configuration keeps the first line
    best2 = fa.autotune(32, L, *ptr, Np, V, w, interlace=False, reps=1)
    keys = sorted((ln.split()[0], ln.split()[1]) for ln in open(w) if not ln.startswith("#"))
    assert keys == [("32", "0"), ("64", "1")], keys
    assert best2["paint_tile"] in (16, 12, 8, 4)


@pytest.mark.parametrize("bk", [16, 32, 128])
def test_bin_block_shapes_agree(torch_cuda, lib, tmp_path, bk):
    import fftgal_amd as fa
    torch = torch_cuda
    Ng, Np, L, Nmu = 64, 100000, 500.0, 6
    _, t = _cat(torch, Np, L, 22)
    ptr = [a.data_ptr() for a in t]
    V = L ** 3
    ref = fa.PowerPlan(Ng, L, Nmu=Nmu, Np_max=Np)
    r0 = ref.run(*ptr, Np, V)
    ref.close()
    w = str(tmp_path / "bk.wsdm")
    _write_wisdom(w, Ng, 1, 16, 0, 0, 0, bk)
    p = fa.PowerPlan(Ng, L, Nmu=Nmu, Np_max=Np, wisdom=w)
    assert p.tuning()["bin_bk"] == bk
    r1 = p.run(*ptr, Np, V)
    p.close()
    _close(r1, r0, V / Np, 1e-12)
    wm = r0["Nmu"] > 0
    assert np.allclose(r1["Pmu"][wm], r0["Pmu"][wm], rtol=1e-10, atol=1e-9 * V / Np)


def test_generic_fft_from_wisdom(torch_cuda, lib, tmp_path):
    """fft_generic = 1 in the wisdom file: generic kernels with a chosen
    launch shape; the result agrees with the specialised kernels' within
    the FP32 FFT error."""
    import fftgal_amd as fa
    torch = torch_cuda
    Ng, Np, L = 96, 200000, 700.0
    _, t = _cat(torch, Np, L, 23)
    ptr = [a.data_ptr() for a in t]
    V = L ** 3
    ref = fa.PowerPlan(Ng, L, Np_max=Np)
    r0 = ref.run(*ptr, Np, V)
    ref.close()
    w = str(tmp_path / "g.wsdm")
    _write_wisdom(w, Ng, 1, 8, 4, 2, 1, 0)
    p = fa.PowerPlan(Ng, L, Np_max=Np, wisdom=w)
    tu = p.tuning()
    assert tu["fft_generic"] == 1 and tu["fft_cols"] == 4 and tu["fft_rows"] == 2
    assert tu["paint_tile"] == 8
    r1 = p.run(*ptr, Np, V)
    p.close()
    sel = (r0["N"] > 0) & (np.arange(len(r0["N"])) < Ng // 2)
    _close(_rows(r1, sel), _rows(r0, sel), V / Np, 1e-5)

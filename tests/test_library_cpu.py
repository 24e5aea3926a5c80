"""CPU checks of the C-ABI library: it loads without a GPU, exports every
function the headers in include/ declare (and nothing else), and its host
code (catalog functions) matches the reference semantics.  No call here
touches the GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(INC, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, re.M):
            names.add(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.check_output(["nm", "-D", "--defined-only", path], text=True)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_headers_declare_the_reference_api():
    names = declared_functions()
    # src/fft.h:27-49, src/gal.h:42-56 (minus the sub-volume selections), src/power.h:7-13
    ref = {"fft_init", "fft_free", "fft_p2g", "fft_g2p", "fft_x2k", "fft_k2x",
           "fft_offset_phase", "fft_deconv", "fft_p2k", "fft_exportf", "fft_importf",
           "gal_init", "gal_resize", "gal_free", "gal_wrap", "gal_rsd", "P", "Pl", "Pmu"}
    assert ref <= names


def test_library_exports_exactly_the_declared_api(lib):
    import fftgal_amd
    names = declared_functions()
    exported = exported_symbols(fftgal_amd.LIBPATH)
    assert names == exported, (names ^ exported)
    for n in names:
        getattr(lib, n)      # resolvable through the loader
    assert set(fftgal_amd.API) == names


def test_reference_struct_layouts(lib):
    import fftgal_amd as fa
    # fft_t: two pointer slots, f, Ng, Ng_pad, Ng3_pad, L, offset[3] (src/fft.h:16-24)
    assert C.sizeof(fa.FftT) == 8 + 8 + 8 + 4 + 4 + 8 + 8 + 24
    assert fa.FftT.f.offset == 16 and fa.FftT.offset.offset == 48
    assert C.sizeof(fa.GalT) == 8 + 8 + 6 * 8 + 8


def test_gal_rsd_and_wrap_match_oracle(lib, orc):
    """gal_rsd (src/gal.c:63-77) and gal_wrap (src/gal.c:51-60) are host code."""
    import fftgal_amd as fa
    rng = np.random.default_rng(2)
    n, L = 1000, 500.0
    x, y, z = rng.random((3, n)) * L
    vx, vy, vz = rng.standard_normal((3, n)) * 300
    cg = fa.Catalog(x, y, z, L ** 3, vx, vy, vz)
    co = orc.Catalog(x, y, z, L ** 3, vx, vy, vz)
    los_g = np.array([1.0, 1.0, 0.0])
    los_o = los_g.copy()
    aH = 0.5 * 100 * np.sqrt(0.29 / 0.125 + 0.71)
    rg = lib.gal_rsd(cg.ptr, los_g.ctypes.data_as(fa.dptr), aH)
    ro = orc.load().oracle_gal_rsd(co.ptr, los_o.ctypes.data_as(orc.oracle.dptr), aH)
    assert np.array_equal(los_g, los_o) and abs(np.linalg.norm(los_g) - 1) < 1e-15
    for name in "xyz":
        a = np.ctypeslib.as_array(getattr(rg.contents, name), shape=(n,))
        b = np.ctypeslib.as_array(getattr(ro.contents, name), shape=(n,))
        assert np.array_equal(a, b)
    assert not rg.contents.vx and not rg.contents.rand
    lib.gal_wrap(rg, L)
    orc.load().oracle_gal_wrap(ro, L)
    a = np.ctypeslib.as_array(rg.contents.z, shape=(n,))
    assert np.all((a >= 0) & (a < L))
    assert np.array_equal(a, np.ctypeslib.as_array(ro.contents.z, shape=(n,)))
    lib.gal_free(rg)
    orc.load().oracle_gal_free(ro)


def test_gal_resize_keeps_prefix(lib):
    import fftgal_amd as fa
    c = fa.Catalog(np.arange(10.0), np.arange(10.0) + 1, np.arange(10.0) + 2, 1.0)
    lib.gal_resize(c.ptr, 4)
    assert c.Np == 4
    assert list(np.ctypeslib.as_array(c.ptr.contents.y, shape=(4,))) == [1.0, 2.0, 3.0, 4.0]
    lib.gal_resize(c.ptr, 1000)
    assert c.Np == 1000


def test_no_gpu_means_clean_error():
    """Without a visible device the fast path reports an error code (the
    drop-in aborts, like the reference's asserts) instead of hanging."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import fftgal_amd as fa\n"
        "try:\n"
        "    fa.PowerPlan(32, 100.0, Np_max=10)\n"
        "except RuntimeError as e:\n"
        "    print('ERR', e)\n" % ROOT)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", CUDA_VISIBLE_DEVICES="-1",
               ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "ERR" in out.stdout and "no HIP device" in out.stdout


def _write_qpm(path, n, seed, trailing_newline=True):
    """A QPM-format file: x y z vx vy vz halo_mass sat_flag halo_id."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * 1000.0
    vel = rng.standard_normal((n, 3)) * 300.0
    mass = 10 ** rng.uniform(11, 15, n)
    sat = rng.integers(0, 2, n)
    hid = rng.integers(0, 10 ** 9, n)
    lines = [f"{p[0]:.6f} {p[1]:.6f} {p[2]:.6f} {v[0]:.4f} {v[1]:.4f} {v[2]:.4f} {m:.6e} {s} {h}"
             for p, v, m, s, h in zip(pos, vel, mass, sat, hid)]
    text = "\n".join(lines) + ("\n" if trailing_newline and lines else "")
    with open(path, "w") as fp:
        fp.write(text)
    return lines


def _gal_arrays(ptr):
    g = ptr.contents
    n = g.Np
    return [np.ctypeslib.as_array(getattr(g, k), shape=(n,)).copy() if n else np.zeros(0)
            for k in ("x", "y", "z", "vx", "vy", "vz")]


@pytest.mark.parametrize("n,trailing,threads,chunk", [(0, True, None, None), (1, True, None, None),
                                                      (1000, False, None, None),
                                                      (120000, True, None, None),
                                                      (120000, True, "1", None),
                                                      (120000, True, "3", None),
                                                      (5000, False, "4", "4096")])
def test_loadqpm_matches_oracle(lib, orc, tmp_path, monkeypatch, n, trailing, threads, chunk):
    """gal_loadqpm_cubic (src/io/qpm.c:9-34), multithreaded over 8 MiB
    chunks, against the oracle's stdio restatement and against Python's
    float() of every field: bit-identical values, the reference's line count
    (a last line without '\\n' is not read), V = L^3."""
    import fftgal_amd as fa
    if threads:
        monkeypatch.setenv("FFTGAL_IO_THREADS", threads)
    if chunk:
        monkeypatch.setenv("FFTGAL_QPM_CHUNK_BYTES", chunk)
    path = str(tmp_path / "mock.qpm")
    lines = _write_qpm(path, n, seed=n, trailing_newline=trailing)
    L = 1000.0
    g = lib.gal_loadqpm_cubic(path.encode(), L)
    o = orc.load().oracle_loadqpm_cubic(path.encode(), L)
    assert o
    nread = n if trailing else n - 1
    assert g.contents.Np == nread == o.contents.Np
    assert g.contents.V == L ** 3
    got, ref = _gal_arrays(g), _gal_arrays(o)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    cols = np.array([[float(v) for v in ln.split()[:6]] for ln in lines[:nread]]).reshape(-1, 6)
    for k in range(6):
        assert np.array_equal(got[k], cols[:, k])
    lib.gal_free(g)
    orc.load().oracle_gal_free(o)


def test_loadqpm_rejects_bad_line(lib, tmp_path):
    """A line that does not start with six numbers aborts with a message, as
    the reference's assert does (run in a child process)."""
    import subprocess
    import sys
    path = tmp_path / "bad.qpm"
    path.write_text("1 2 3 4 5 6 7 0 1\n1 2 three 4 5 6 7 0 1\n")
    code = ("import fftgal_amd as fa; lib = fa.load(); "
            f"lib.gal_loadqpm_cubic({str(path)!r}.encode(), 10.0)")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(INC), timeout=120)
    assert r.returncode != 0
    assert "line 2" in r.stderr


def test_fft_init_unwritable_wisdom_aborts(lib, tmp_path):
    """The reference asserts on a wisdom file it cannot export
    (src/fft.c:40); fft_init checks that it can write the file before any
    device work, so the abort shows without a GPU (run in a child process)."""
    import subprocess
    import sys
    bad = tmp_path / "no_such_dir" / "64.wsdm"
    code = ("import fftgal_amd as fa; fa.Grid(64, 10.0, wisdom=%r); print('returned')" % str(bad))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(INC), timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "cannot write the wisdom file" in r.stderr
    assert not bad.parent.exists()

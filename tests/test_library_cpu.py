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

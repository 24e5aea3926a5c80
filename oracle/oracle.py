"""ctypes view of oracle/liboracle.so (FP64 CPU restatement of the reference).

TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Each wrapper names
the reference function it mirrors (eelregit/fftgal src/fft.c, src/power.c,
src/gal.c).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dptr = C.POINTER(C.c_double)
lptr = C.POINTER(C.c_long)


class OfftPlan(C.Structure):
    _fields_ = [("n", C.c_int), ("nstage", C.c_int), ("radix", C.c_int * 64),
                ("tw", dptr)]


class Offt(C.Structure):
    _fields_ = [("f", dptr), ("Ng", C.c_int), ("Ng_pad", C.c_int),
                ("Ng3_pad", C.c_long), ("L", C.c_double),
                ("offset", C.c_double * 3),
                ("pN", C.POINTER(OfftPlan)), ("pH", C.POINTER(OfftPlan))]


class Ogal(C.Structure):
    pass


Ogal._fields_ = [("Np", C.c_long), ("V", C.c_double),
                 ("x", dptr), ("y", dptr), ("z", dptr),
                 ("vx", dptr), ("vy", dptr), ("vz", dptr),
                 ("rand", C.POINTER(Ogal))]


def load():
    """Load (building if needed) the oracle shared library."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # FFTGAL_ORACLE_LIB selects another build (the sanitizer build of
    # oracle/Makefile's `asan` target)
    path = os.environ.get("FFTGAL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    lib = C.CDLL(path)
    G = C.POINTER(Offt)
    P = C.POINTER(Ogal)
    d3 = C.c_double * 3
    sig = {
        "oracle_fft_init": (G, [C.c_int, C.c_double]),
        "oracle_fft_free": (None, [G]),
        "oracle_fft_p2g": (None, [G, P, dptr]),
        "oracle_fft_x2k": (None, [G, C.c_int]),
        "oracle_fft_k2x": (None, [G, C.c_int]),
        "oracle_fft_offset_phase": (None, [G, C.c_int]),
        "oracle_fft_deconv": (None, [G]),
        "oracle_fft_p2k": (None, [G, P]),
        "oracle_fft_exportf": (dptr, [G]),
        "oracle_fft_importf": (None, [G, dptr]),
        "oracle_P": (None, [G, P]),
        "oracle_NK": (C.c_int, [C.c_int, C.c_double, C.c_double]),
        "oracle_Pl_bins": (C.c_int, [G, dptr, C.c_double, dptr, lptr, dptr]),
        "oracle_Pmu_bins": (C.c_int, [G, dptr, C.c_double, C.c_int, dptr, lptr, dptr]),
        "oracle_Pl": (None, [G, P, dptr, C.c_double, C.c_char_p]),
        "oracle_Pmu": (None, [G, P, dptr, C.c_double, C.c_char_p]),
        "oracle_hat": (None, [dptr]),
        "oracle_gal_init": (P, [C.c_long, C.c_double]),
        "oracle_gal_free": (None, [P]),
        "oracle_gal_wrap": (None, [P, C.c_double]),
        "oracle_gal_rsd": (P, [P, dptr, C.c_double]),
        "oracle_loadqpm_cubic": (P, [C.c_char_p, C.c_double]),
        "oracle_dl_filter": (None, [G, C.c_int, C.c_int, dptr, C.c_int, C.c_double, dptr]),
        "oracle_gal_subbox": (P, [P, dptr, C.c_double]),
        "oracle_gal_subsphere": (P, [P, dptr, C.c_double]),
        "oracle_subvolume_deconv": (None, [G, C.c_int, C.c_int]),
        "oracle_set_threads": (None, [C.c_int]),
        "oracle_threads": (C.c_int, []),
        "offt64_r2c_3d": (None, [G]),
        "offt64_c2r_3d": (None, [G]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    del d3
    _LIB = lib
    return lib


def set_threads(n):
    """Threads for the large sweeps (paint by i-slabs in the serial particle
    order, whole 1D FFTs, k-space planes); results are bit-identical to one
    thread.  The bin sums stay serial."""
    load().oracle_set_threads(int(n))


def default_threads():
    """CPU threads this process may use, at most 16."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _dp(a):
    return a.ctypes.data_as(dptr)


def hat(los):
    """src/gal.h:19-26 (returns the normalised copy)."""
    v = np.array(los, dtype=np.float64)
    load().oracle_hat(_dp(v))
    return v


def NK(Ng, L, dK):
    """Number of k bins, src/power.c:36."""
    return load().oracle_NK(Ng, L, dK)


class Catalog:
    """gal_t (src/gal.h:29-39) holding FP64 copies of the given positions."""

    def __init__(self, x, y, z, V, vx=None, vy=None, vz=None, rand=None):
        lib = load()
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.shape[0]
        self.ptr = lib.oracle_gal_init(n, float(V))
        g = self.ptr.contents
        for name, arr in (("x", x), ("y", y), ("z", z)):
            a = np.ascontiguousarray(arr, dtype=np.float64)
            if n:
                C.memmove(getattr(g, name), a.ctypes.data, n * 8)
        if vx is not None:
            self._v = [np.ascontiguousarray(v, dtype=np.float64) for v in (vx, vy, vz)]
            # malloc'd copies so that oracle_gal_free can release them
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            libc.malloc.argtypes = [C.c_size_t]
            for name, a in zip(("vx", "vy", "vz"), self._v):
                p = libc.malloc(max(n, 1) * 8)
                C.memmove(p, a.ctypes.data, n * 8)
                setattr(g, name, C.cast(p, dptr))
        self.rand = rand
        if rand is not None:
            g.rand = rand.ptr
            rand._owned_by_parent = True
        self._owned_by_parent = False

    @property
    def Np(self):
        return self.ptr.contents.Np

    def free(self):
        if self.ptr and not self._owned_by_parent:
            load().oracle_gal_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """fft_t (src/fft.h:16-24) backed by the oracle."""

    def __init__(self, Ng, L):
        self.lib = load()
        self.ptr = self.lib.oracle_fft_init(int(Ng), float(L))
        g = self.ptr.contents
        self.Ng, self.Ng_pad, self.L = g.Ng, g.Ng_pad, g.L
        buf = (C.c_double * g.Ng3_pad).from_address(C.addressof(g.f.contents))
        self.f = np.frombuffer(buf, dtype=np.float64).reshape(self.Ng, self.Ng, self.Ng_pad)

    @property
    def offset(self):
        return list(self.ptr.contents.offset)

    def fk(self):
        """Complex view [Ng, Ng, Ng/2+1] of the half spectrum (copy)."""
        h = self.Ng // 2 + 1
        v = self.f[:, :, : 2 * h]
        return v[..., 0::2] + 1j * v[..., 1::2]

    def p2g(self, cat, offset=(0.0, 0.0, 0.0)):
        """src/fft.c:57-131; returns the reduced offset like the reference."""
        o = np.array(offset, dtype=np.float64)
        self.lib.oracle_fft_p2g(self.ptr, cat.ptr, _dp(o))
        return o

    def x2k(self, phase_on):
        self.lib.oracle_fft_x2k(self.ptr, int(phase_on))

    def k2x(self, phase_off):
        self.lib.oracle_fft_k2x(self.ptr, int(phase_off))

    def offset_phase(self, onoff):
        self.lib.oracle_fft_offset_phase(self.ptr, int(onoff))

    def deconv(self):
        self.lib.oracle_fft_deconv(self.ptr)

    def p2k(self, cat):
        self.lib.oracle_fft_p2k(self.ptr, cat.ptr)

    def P(self, cat):
        self.lib.oracle_P(self.ptr, cat.ptr)

    def Pl_bins(self, los, dK):
        nk = NK(self.Ng, self.L, dK)
        K = np.zeros(nk)
        N = np.zeros(nk, dtype=np.int64)
        P = np.zeros((nk, 4))
        l = np.array(los, dtype=np.float64)
        self.lib.oracle_Pl_bins(self.ptr, _dp(l), float(dK), _dp(K),
                                N.ctypes.data_as(lptr), _dp(P))
        return K, N, P

    def Pmu_bins(self, los, dK, Nmu):
        nk = NK(self.Ng, self.L, dK)
        K = np.zeros((nk, Nmu))
        N = np.zeros((nk, Nmu), dtype=np.int64)
        P = np.zeros((nk, Nmu))
        l = np.array(los, dtype=np.float64)
        self.lib.oracle_Pmu_bins(self.ptr, _dp(l), float(dK), int(Nmu), _dp(K),
                                 N.ctypes.data_as(lptr), _dp(P))
        return K, N, P

    def Pl(self, cat, los, dK, path):
        l = np.array(los, dtype=np.float64)
        self.lib.oracle_Pl(self.ptr, cat.ptr, _dp(l), float(dK), path.encode())
        return l

    def Pmu(self, cat, los, dK, path):
        l = np.array(los, dtype=np.float64)
        self.lib.oracle_Pmu(self.ptr, cat.ptr, _dp(l), float(dK), path.encode())
        return l

    def free(self):
        if self.ptr:
            self.lib.oracle_fft_free(self.ptr)
        self.ptr = None
        self.f = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fft_r2c_3d(x):
    """Unnormalised forward 3D r2c of a cubic real array with the oracle FFT."""
    Ng = x.shape[0]
    g = Grid(Ng, 1.0)
    g.f[:, :, :Ng] = x
    g.lib.offt64_r2c_3d(g.ptr)
    out = g.fk()
    g.free()
    return out


def fft_c2r_3d(X, Ng):
    """Unnormalised inverse 3D c2r (FFTW convention) with the oracle FFT."""
    g = Grid(Ng, 1.0)
    h = Ng // 2 + 1
    g.f[:, :, 0:2 * h:2] = X.real
    g.f[:, :, 1:2 * h:2] = X.imag
    g.lib.offt64_c2r_3d(g.ptr)
    out = g.f[:, :, :Ng].copy()
    g.free()
    return out


def power_pipeline(x, y, z, V, Ng, L, interlace=True, los=(0, 0, 1), dK=None,
                   Nmu=0, timings=None):
    """Reference call order for one catalog (src/app/P.c:53-66):
    fft_p2k (or p2g + x2k + deconv without interlacing) -> P -> Pl bins
    (-> Pmu bins).  Returns dict with K, N, Pl (NK x 4) and optional wedges."""
    import time
    t0 = time.perf_counter()
    cat = Catalog(x, y, z, V)
    g = Grid(Ng, L)
    t1 = time.perf_counter()
    if interlace:
        g.p2k(cat)
    else:
        g.p2g(cat, (0.0, 0.0, 0.0))
        g.x2k(0)
        g.deconv()
    t2 = time.perf_counter()
    g.P(cat)
    if dK is None:
        dK = 2 * np.pi / L
    K, N, P = g.Pl_bins(los, dK)
    out = {"K": K, "N": N, "Pl": P, "dK": dK}
    if Nmu:
        Km, Nm, Pm = g.Pmu_bins(los, dK, Nmu)
        out.update(Kmu=Km, Nmu=Nm, Pmu=Pm)
    t3 = time.perf_counter()
    if timings is not None:
        timings.update(setup=t1 - t0, p2k=t2 - t1, bins=t3 - t2)
    g.free()
    cat.free()
    return out

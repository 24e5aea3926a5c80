"""fftgal_amd -- the fftgal power-spectrum path on one MI355X.

The product is the C shared library ``fftgal_amd/libfftgal_amd.so`` with the
reference's C API (include/fft.h, gal.h, power.h) plus the additive
include/fftgal_ext.h and include/fftgal_batch.h.  This module is a thin
ctypes view of that library for Python callers, tests and the benchmark; it
computes nothing itself and raises if the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FFTGAL_LIB selects another build of the same library (the host-sanitizer
# build of fftgal_amd/csrc/Makefile's `asan` target)
LIBPATH = os.environ.get("FFTGAL_LIB") or os.path.join(_HERE, "libfftgal_amd.so")
_LIB = None

dptr = C.POINTER(C.c_double)
lptr = C.POINTER(C.c_long)
fptr = C.c_void_p  # device float32 arrays are passed as raw addresses


class FftT(C.Structure):
    """fft_t of include/fft.h (same fields as the reference's src/fft.h:16-24)."""
    _fields_ = [("x2k", C.c_void_p), ("k2x", C.c_void_p), ("f", dptr),
                ("Ng", C.c_int), ("Ng_pad", C.c_int), ("Ng3_pad", C.c_long),
                ("L", C.c_double), ("offset", C.c_double * 3)]


class GalT(C.Structure):
    """gal_t of include/gal.h (src/gal.h:29-39)."""


GalT._fields_ = [("Np", C.c_long), ("V", C.c_double),
                 ("x", dptr), ("y", dptr), ("z", dptr),
                 ("vx", dptr), ("vy", dptr), ("vz", dptr),
                 ("rand", C.POINTER(GalT))]


class PowerDesc(C.Structure):
    """fftgal_power_desc of include/fftgal_batch.h."""
    _fields_ = [("Ng", C.c_int), ("L", C.c_double), ("interlace", C.c_int),
                ("dK", C.c_double), ("los", C.c_double * 3), ("Nmu", C.c_int),
                ("Np_max", C.c_long), ("paint_tile", C.c_int),
                ("wisdom", C.c_char_p)]


# every C entry point the headers in include/ declare: name -> (restype, argtypes)
API = {
    # include/fft.h
    "fft_init": (C.POINTER(FftT), [C.c_int, C.c_double, C.c_char_p]),
    "fft_free": (None, [C.POINTER(FftT)]),
    "fft_p2g": (None, [C.POINTER(FftT), C.POINTER(GalT), dptr]),
    "fft_g2p": (dptr, [C.POINTER(FftT), C.POINTER(GalT)]),
    "fft_x2k": (None, [C.POINTER(FftT), C.c_int]),
    "fft_k2x": (None, [C.POINTER(FftT), C.c_int]),
    "fft_offset_phase": (None, [C.POINTER(FftT), C.c_int]),
    "fft_deconv": (None, [C.POINTER(FftT)]),
    "fft_p2k": (None, [C.POINTER(FftT), C.POINTER(GalT)]),
    "fft_exportf": (dptr, [C.POINTER(FftT)]),
    "fft_importf": (None, [C.POINTER(FftT), dptr]),
    # include/gal.h
    "gal_init": (C.POINTER(GalT), [C.c_long, C.c_double]),
    "gal_resize": (None, [C.POINTER(GalT), C.c_long]),
    "gal_free": (None, [C.POINTER(GalT)]),
    "gal_wrap": (None, [C.POINTER(GalT), C.c_double]),
    "gal_rsd": (C.POINTER(GalT), [C.POINTER(GalT), dptr, C.c_double]),
    # include/power.h
    "P": (None, [C.POINTER(FftT), C.POINTER(GalT)]),
    "Pl": (None, [C.POINTER(FftT), C.POINTER(GalT), dptr, C.c_double, C.c_char_p]),
    "Pmu": (None, [C.POINTER(FftT), C.POINTER(GalT), dptr, C.c_double, C.c_char_p]),
    # include/fftgal_ext.h
    "fft_set_device_resident": (None, [C.POINTER(FftT), C.c_int]),
    "fft_sync_host": (None, [C.POINTER(FftT)]),
    "fft_device_grid": (C.c_void_p, [C.POINTER(FftT)]),
    "fft_p2g_device": (None, [C.POINTER(FftT), C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_long, C.c_double, dptr]),
    "fft_p2k_device": (None, [C.POINTER(FftT), C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_long, C.c_double]),
    "P_params": (None, [C.POINTER(FftT), C.c_long, C.c_double, C.c_long]),
    "fftgal_NK": (C.c_int, [C.c_int, C.c_double, C.c_double]),
    "Pl_bins": (C.c_int, [C.POINTER(FftT), dptr, C.c_double, dptr, lptr, dptr]),
    "Pmu_bins": (C.c_int, [C.POINTER(FftT), dptr, C.c_double, C.c_int, dptr, lptr, dptr]),
    "Pl_bins_noalloc": (C.c_int, [C.POINTER(FftT), dptr, C.c_double, dptr, lptr, dptr]),
    # include/fftgal_batch.h
    "fftgal_last_error": (C.c_char_p, []),
    "fftgal_power_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PowerDesc)]),
    "fftgal_power_destroy": (C.c_int, [C.c_void_p]),
    "fftgal_power_NK": (C.c_int, [C.c_void_p]),
    "fftgal_power_enqueue": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_long, C.c_double]),
    "fftgal_power_sync": (C.c_int, [C.c_void_p]),
    "fftgal_power_results": (C.c_int, [C.c_void_p, dptr, lptr, dptr, dptr, lptr, dptr]),
    "fftgal_power_paint_only": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_long, C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p)]),
    "fftgal_power_fixed_bits": (C.c_int, [C.c_long]),
    "fftgal_power_fixed_scale": (C.c_double, []),
    "fftgal_power_grid_sets": (C.c_int, [C.c_void_p]),
    "fftgal_power_spectra": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p)]),
    "fftgal_power_tuning": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "fftgal_power_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_long, C.c_double, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fftgal_power_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "fftgal_power_fft_generic": (C.c_int, [C.c_void_p, C.c_int]),
    "fftgal_power_stream": (C.c_void_p, [C.c_void_p]),
    "fftgal_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_ulong]),
    "fftgal_power_stages": (C.c_int, []),
    "fftgal_power_stage_name": (C.c_char_p, [C.c_int]),
    "fftgal_power_profile": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_long, C.c_double, C.POINTER(C.c_float)]),
    "fft_dl_filter": (None, [C.POINTER(FftT), C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int,
                             C.c_double, C.POINTER(C.c_double)]),
    "gal_subbox": (C.POINTER(GalT), [C.POINTER(GalT), C.POINTER(C.c_double), C.c_double]),
    "gal_subsphere": (C.POINTER(GalT), [C.POINTER(GalT), C.POINTER(C.c_double), C.c_double]),
    "fft_subvolume_deconv": (None, [C.POINTER(FftT), C.c_int, C.c_int]),
    # include/io.h
    "gal_loadqpm_cubic": (C.POINTER(GalT), [C.c_char_p, C.c_double]),
    "fftgal_qpm_load_device": (C.c_int, [C.c_char_p, C.c_long, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_long)]),
    "fftgal_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_void_p,
                                      C.POINTER(C.c_double), C.c_int, C.c_char_p]),
    "fftgal_batch_destroy": (C.c_int, [C.c_void_p]),
    "fftgal_batch_NK": (C.c_int, [C.c_void_p, C.c_int]),
    "fftgal_batch_enqueue": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "fftgal_batch_sync": (C.c_int, [C.c_void_p]),
    "fftgal_batch_results": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_long), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_long),
                                       C.POINTER(C.c_double)]),
    "fftgal_batch_profile": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "fftgal_batch_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "fftgal_batch_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "fftgal_batch_grid_sets": (C.c_int, [C.c_void_p]),
    "fftgal_batch_rounds": (C.c_int, [C.c_void_p]),
    "fftgal_batch_pool_bytes": (C.c_size_t, [C.c_void_p]),
    "fftgal_batch_stream": (C.c_void_p, [C.c_void_p]),
}


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libfftgal_amd.so (built by __graft_entry__.build() or
    `make -C fftgal_amd/csrc`).  Raises LibraryMissing if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise LibraryMissing(
            f"{LIBPATH} not found: build it with `make -C fftgal_amd/csrc` "
            "(there is no fallback implementation)")
    lib = C.CDLL(LIBPATH)
    for name, (res, args) in API.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(dptr)


def last_error():
    return load().fftgal_last_error().decode()


class Catalog:
    """A gal_t owned by the library (gal_init / gal_free)."""

    def __init__(self, x, y, z, V, vx=None, vy=None, vz=None):
        lib = load()
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = int(x.shape[0])
        self.ptr = lib.gal_init(n, float(V))
        g = self.ptr.contents
        for name, arr in (("x", x), ("y", y), ("z", z)):
            a = np.ascontiguousarray(arr, dtype=np.float64)
            if n:
                C.memmove(getattr(g, name), a.ctypes.data, n * 8)
        if vx is not None:
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            libc.malloc.argtypes = [C.c_size_t]
            for name, arr in (("vx", vx), ("vy", vy), ("vz", vz)):
                a = np.ascontiguousarray(arr, dtype=np.float64)
                p = libc.malloc(max(n, 1) * 8)
                C.memmove(p, a.ctypes.data, n * 8)
                setattr(g, name, C.cast(p, dptr))

    @property
    def Np(self):
        return self.ptr.contents.Np

    def free(self):
        if getattr(self, "ptr", None):
            load().gal_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """fft_t of the drop-in API; ``f`` is a numpy view of the host grid."""

    def __init__(self, Ng, L, wisdom="FFTW_ESTIMATE", resident=False):
        self.lib = load()
        self.ptr = self.lib.fft_init(int(Ng), float(L), wisdom.encode())
        g = self.ptr.contents
        self.Ng, self.Ng_pad, self.L = g.Ng, g.Ng_pad, g.L
        buf = (C.c_double * g.Ng3_pad).from_address(C.addressof(g.f.contents))
        self.f = np.frombuffer(buf, dtype=np.float64).reshape(self.Ng, self.Ng, self.Ng_pad)
        if resident:
            self.lib.fft_set_device_resident(self.ptr, 1)

    @property
    def offset(self):
        return list(self.ptr.contents.offset)

    def fk(self):
        h = self.Ng // 2 + 1
        v = self.f[:, :, :2 * h]
        return v[..., 0::2] + 1j * v[..., 1::2]

    def p2g(self, cat, offset=(0.0, 0.0, 0.0)):
        o = np.array(offset, dtype=np.float64)
        self.lib.fft_p2g(self.ptr, cat.ptr, _dp(o))
        return o

    def x2k(self, phase_on):
        self.lib.fft_x2k(self.ptr, int(phase_on))

    def k2x(self, phase_off):
        self.lib.fft_k2x(self.ptr, int(phase_off))

    def offset_phase(self, onoff):
        self.lib.fft_offset_phase(self.ptr, int(onoff))

    def deconv(self):
        self.lib.fft_deconv(self.ptr)

    def p2k(self, cat):
        self.lib.fft_p2k(self.ptr, cat.ptr)

    def P(self, cat):
        self.lib.P(self.ptr, cat.ptr)

    def sync_host(self):
        self.lib.fft_sync_host(self.ptr)

    def Pl_bins(self, los, dK):
        nk = self.lib.fftgal_NK(self.Ng, self.L, float(dK))
        K = np.zeros(nk)
        N = np.zeros(nk, dtype=np.int64)
        P = np.zeros((nk, 4))
        l = np.array(los, dtype=np.float64)
        self.lib.Pl_bins(self.ptr, _dp(l), float(dK), _dp(K), N.ctypes.data_as(lptr), _dp(P))
        return K, N, P

    def Pmu_bins(self, los, dK, Nmu):
        nk = self.lib.fftgal_NK(self.Ng, self.L, float(dK))
        K = np.zeros((nk, Nmu))
        N = np.zeros((nk, Nmu), dtype=np.int64)
        P = np.zeros((nk, Nmu))
        l = np.array(los, dtype=np.float64)
        self.lib.Pmu_bins(self.ptr, _dp(l), float(dK), int(Nmu), _dp(K),
                          N.ctypes.data_as(lptr), _dp(P))
        return K, N, P

    def Pl(self, cat, los, dK, path):
        l = np.array(los, dtype=np.float64)
        self.lib.Pl(self.ptr, cat.ptr, _dp(l), float(dK), path.encode())
        return l

    def Pmu(self, cat, los, dK, path):
        l = np.array(los, dtype=np.float64)
        self.lib.Pmu(self.ptr, cat.ptr, _dp(l), float(dK), path.encode())
        return l

    def free(self):
        if getattr(self, "ptr", None):
            self.lib.fft_free(self.ptr)
        self.ptr = None
        self.f = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PowerPlan:
    """fftgal_power_t: the mixed-precision pipeline for one geometry.

    Positions are device float32 arrays given by address (e.g. a torch CUDA
    tensor's data_ptr()); the caller makes sure they are written before
    enqueue (e.g. torch.cuda.synchronize())."""

    def __init__(self, Ng, L, interlace=True, dK=None, los=(0.0, 0.0, 1.0), Nmu=0,
                 Np_max=1, paint_tile=0, wisdom=None):
        self.lib = load()
        d = PowerDesc()
        d.Ng = int(Ng)
        d.L = float(L)
        d.interlace = 1 if interlace else 0
        d.dK = float(dK if dK is not None else 2 * np.pi / L)
        d.los = (C.c_double * 3)(*[float(v) for v in los])
        d.Nmu = int(Nmu)
        d.Np_max = int(Np_max)
        d.paint_tile = int(paint_tile)
        d.wisdom = wisdom.encode() if wisdom else None
        self.desc = d
        h = C.c_void_p()
        rc = self.lib.fftgal_power_create(C.byref(h), C.byref(d))
        if rc:
            raise RuntimeError(f"fftgal_power_create failed ({rc}): {last_error()}")
        self.h = h
        self.NK = self.lib.fftgal_power_NK(h)
        self.Nmu = int(Nmu)
        self.Ng = int(Ng)

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {last_error()}")

    def enqueue(self, x, y, z, Np, V):
        self._check(self.lib.fftgal_power_enqueue(self.h, x, y, z, int(Np), float(V)),
                    "fftgal_power_enqueue")

    def sync(self):
        self._check(self.lib.fftgal_power_sync(self.h), "fftgal_power_sync")

    def results(self):
        nk, nm = self.NK, max(self.Nmu, 1)
        K = np.zeros(nk)
        N = np.zeros(nk, dtype=np.int64)
        P = np.zeros((nk, 4))
        Km = np.zeros((nk, nm))
        Nm = np.zeros((nk, nm), dtype=np.int64)
        Pm = np.zeros((nk, nm))
        self._check(self.lib.fftgal_power_results(self.h, _dp(K), N.ctypes.data_as(lptr), _dp(P),
                                                  _dp(Km), Nm.ctypes.data_as(lptr), _dp(Pm)),
                    "fftgal_power_results")
        out = {"K": K, "N": N, "Pl": P}
        if self.Nmu:
            out.update(Kmu=Km, Nmu=Nm, Pmu=Pm)
        return out

    def run(self, x, y, z, Np, V):
        self.enqueue(x, y, z, Np, V)
        self.sync()
        return self.results()

    def profile(self, x, y, z, Np, V):
        """Eager run with events between stages: {stage name: ms}."""
        n = self.lib.fftgal_power_stages()
        ms = (C.c_float * n)()
        self._check(self.lib.fftgal_power_profile(self.h, x, y, z, int(Np), float(V), ms),
                    "fftgal_power_profile")
        return {self.lib.fftgal_power_stage_name(i).decode(): float(ms[i]) for i in range(n)}

    def paint_only(self, x, y, z, Np):
        """Run only the paint stage; returns the int64 grids (plain, dual or
        None) copied to host as [Ng, Ng, Ng] arrays in units of fixed_scale()
        particles."""
        a, b = C.c_void_p(), C.c_void_p()
        self._check(self.lib.fftgal_power_paint_only(self.h, x, y, z, int(Np), C.byref(a),
                                                     C.byref(b)), "fftgal_power_paint_only")
        out = []
        for ptr in (a.value, b.value):
            if not ptr:
                out.append(None)
                continue
            h = np.empty((self.Ng,) * 3, dtype=np.int64)
            self._check(self.lib.fftgal_memcpy_d2h(h.ctypes.data, ptr, h.nbytes),
                        "fftgal_memcpy_d2h")
            out.append(h)
        return tuple(out)

    def fixed_bits(self, n):
        """fraction bits of a paint tile whose work items hold n particles"""
        return self.lib.fftgal_power_fixed_bits(int(n))

    def fixed_scale(self):
        """particle mass of one int64 grid unit (about 2^-25)"""
        return self.lib.fftgal_power_fixed_scale()

    def grid_sets(self):
        """Number of int64 grid sets (2: catalogs alternate and each paint
        clears the other set)."""
        return self.lib.fftgal_power_grid_sets(self.h)

    def spectra(self):
        a, b = C.c_void_p(), C.c_void_p()
        self.lib.fftgal_power_spectra(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def tuning(self):
        t, c, r, g, k = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.lib.fftgal_power_tuning(self.h, C.byref(t), C.byref(c), C.byref(r), C.byref(g),
                                     C.byref(k))
        return {"paint_tile": t.value, "fft_cols": c.value, "fft_rows": r.value,
                "fft_generic": g.value, "bin_bk": k.value}

    def use_graph(self, on):
        self.lib.fftgal_power_use_graph(self.h, 1 if on else 0)

    def fft_generic(self, on):
        """Force the generic FFT kernels (True) or restore the default;
        returns True if the generic kernels are in use."""
        r = self.lib.fftgal_power_fft_generic(self.h, 1 if on else 0)
        self._check(min(r, 0), "fftgal_power_fft_generic")
        return bool(r)

    def spectra_host(self):
        """The FP32 spectra of the last run copied to host as complex64
        arrays [Ng, Ng, Ng//2] in the packed half-spectrum layout."""
        out = []
        for ptr in self.spectra():
            if not ptr:
                out.append(None)
                continue
            h = np.empty((self.Ng, self.Ng, self.Ng // 2), np.complex64)
            self._check(self.lib.fftgal_memcpy_d2h(h.ctypes.data, ptr, h.nbytes),
                        "fftgal_memcpy_d2h")
            out.append(h)
        return tuple(out)

    def stream(self):
        return self.lib.fftgal_power_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fftgal_power_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(Ng, L, x, y, z, Np, V, wisdom, interlace=True, dK=None, los=(0.0, 0.0, 1.0),
             Nmu=0, reps=2):
    """fftgal_power_autotune: measure the paint tile, FFT kernels and bin
    block on this catalog (device float32 positions by address), store the
    fastest in the wisdom file and return it."""
    lib = load()
    d = PowerDesc()
    d.Ng = int(Ng)
    d.L = float(L)
    d.interlace = 1 if interlace else 0
    d.dK = float(dK if dK is not None else 2 * np.pi / L)
    d.los = (C.c_double * 3)(*[float(v) for v in los])
    d.Nmu = int(Nmu)
    d.Np_max = int(Np)
    d.paint_tile = 0
    d.wisdom = wisdom.encode()
    t, g, k = C.c_int(), C.c_int(), C.c_int()
    rc = lib.fftgal_power_autotune(C.byref(d), x, y, z, int(Np), float(V), int(reps),
                                   C.byref(t), C.byref(g), C.byref(k))
    if rc:
        raise RuntimeError(f"fftgal_power_autotune failed ({rc}): {last_error()}")
    return {"paint_tile": t.value, "fft_generic": g.value, "bin_bk": k.value}


class JobDesc(C.Structure):
    """fftgal_job_desc (include/fftgal_batch.h)."""
    _fields_ = [("Ng", C.c_int), ("L", C.c_double), ("interlace", C.c_int),
                ("dK", C.c_double), ("Np_max", C.c_long), ("paint_tile", C.c_int)]


class BatchPlan:
    """fftgal_batch_t: many catalogs, each on its own grid, through one
    launch per pipeline stage.  ``jobs`` is a list of dicts with keys Ng, L
    (grid box = catalog box / fold), Np_max and optionally interlace (True),
    dK (2 pi / L) and paint_tile (0)."""

    def __init__(self, jobs, los=(0.0, 0.0, 1.0), Nmu=0, wisdom=None):
        self.lib = load()
        n = len(jobs)
        arr = (JobDesc * n)()
        for i, j in enumerate(jobs):
            arr[i].Ng = int(j["Ng"])
            arr[i].L = float(j["L"])
            arr[i].interlace = 1 if j.get("interlace", True) else 0
            arr[i].dK = float(j.get("dK", 2 * np.pi / j["L"]))
            arr[i].Np_max = int(j["Np_max"])
            arr[i].paint_tile = int(j.get("paint_tile", 0))
        self.n = n
        self.Nmu = int(Nmu)
        lo = (C.c_double * 3)(*los)
        h = C.c_void_p()
        rc = self.lib.fftgal_batch_create(C.byref(h), n, C.cast(arr, C.c_void_p), lo, int(Nmu),
                                          wisdom.encode() if wisdom else None)
        if rc:
            raise RuntimeError(f"fftgal_batch_create failed ({rc}): {last_error()}")
        self.h = h
        self.NK = [self.lib.fftgal_batch_NK(h, i) for i in range(n)]

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {last_error()}")

    def _args(self, cats):
        """cats: list of (x_ptr, y_ptr, z_ptr, Np, V) per catalog."""
        n = self.n
        if len(cats) != n:
            raise ValueError(f"expected {n} catalogs, got {len(cats)}")
        xs = (C.c_void_p * n)(*[c[0] for c in cats])
        ys = (C.c_void_p * n)(*[c[1] for c in cats])
        zs = (C.c_void_p * n)(*[c[2] for c in cats])
        nps = (C.c_long * n)(*[int(c[3]) for c in cats])
        vs = (C.c_double * n)(*[float(c[4]) for c in cats])
        self._keep = (xs, ys, zs, nps, vs)
        return xs, ys, zs, nps, vs

    def enqueue(self, cats):
        self._check(self.lib.fftgal_batch_enqueue(self.h, *self._args(cats)),
                    "fftgal_batch_enqueue")

    def sync(self):
        self._check(self.lib.fftgal_batch_sync(self.h), "fftgal_batch_sync")

    def results(self, job):
        nk, nm = self.NK[job], max(self.Nmu, 1)
        K = np.zeros(nk)
        N = np.zeros(nk, dtype=np.int64)
        P = np.zeros((nk, 4))
        Km = np.zeros((nk, nm))
        Nm = np.zeros((nk, nm), dtype=np.int64)
        Pm = np.zeros((nk, nm))
        self._check(self.lib.fftgal_batch_results(self.h, int(job), _dp(K),
                                                  N.ctypes.data_as(lptr), _dp(P), _dp(Km),
                                                  Nm.ctypes.data_as(lptr), _dp(Pm)),
                    "fftgal_batch_results")
        out = {"K": K, "N": N, "Pl": P}
        if self.Nmu:
            out.update(Kmu=Km, Nmu=Nm, Pmu=Pm)
        return out

    def run(self, cats):
        self.enqueue(cats)
        self.sync()
        return [self.results(j) for j in range(self.n)]

    def profile(self, cats):
        """Eager run with events between stages: {stage name: ms} (each
        stage covers every catalog)."""
        n = self.lib.fftgal_power_stages()
        ms = (C.c_float * n)()
        self._check(self.lib.fftgal_batch_profile(self.h, *self._args(cats), ms),
                    "fftgal_batch_profile")
        return {self.lib.fftgal_power_stage_name(i).decode(): float(ms[i]) for i in range(n)}

    def use_graph(self, on):
        self.lib.fftgal_batch_use_graph(self.h, 1 if on else 0)

    def tile(self, job):
        return self.lib.fftgal_batch_tile(self.h, int(job))

    def grid_sets(self):
        return self.lib.fftgal_batch_grid_sets(self.h)

    def rounds(self):
        """rounds the catalogs run in (the pool bound, FFTGAL_BATCH_BYTES)"""
        return self.lib.fftgal_batch_rounds(self.h)

    def pool_bytes(self):
        return self.lib.fftgal_batch_pool_bytes(self.h)

    def stream(self):
        return self.lib.fftgal_batch_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fftgal_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

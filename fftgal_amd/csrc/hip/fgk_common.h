// Shared device-side helpers for the gfx950 kernels of libfftgal_amd.
#ifndef FGK_COMMON_H
#define FGK_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

int fgk_hip_fail(hipError_t e, const char *what, const char *file, int line);

#define FGK_HIP(call)                                                        \
    do {                                                                     \
        hipError_t fgk_e_ = (call);                                          \
        if (fgk_e_ != hipSuccess)                                            \
            return fgk_hip_fail(fgk_e_, #call, __FILE__, __LINE__);          \
    } while (0)

#define FGK_LAUNCHED() FGK_HIP(hipGetLastError())

// complex value in registers / LDS
template <typename T> struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return cmk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return cmk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return cmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return cmk(a.x * s, a.y * s); }
// multiply by sign*i
template <typename T>
__device__ __forceinline__ cpx<T> cmul_si(cpx<T> a, int sign) {
    return sign < 0 ? cmk(a.y, -a.x) : cmk(-a.y, a.x);
}

static inline unsigned fgk_ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2.  xcd_block maps blockIdx to a logical block such that each
// XCD owns a contiguous range of logical blocks: neighbouring blocks share
// cache lines at their edges, and this keeps those lines in one XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    unsigned q = nb >> 3;
    if (b < (q << 3)) return (b & 7u) * q + (b >> 3);
    return b;
}


#endif

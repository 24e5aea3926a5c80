// Device runtime for libfftgal_amd: device selection, memory, streams,
// events, hipGraph capture and error reporting behind the C-ABI in fgk.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "../fgk.h"
#include "fgk_common.h"

static thread_local char g_err[1024];
static int g_inited = 0;

extern "C" const char *fgk_last_error(void) { return g_err; }

extern "C" void fgk_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fgk_hip_fail(hipError_t e, const char *what, const char *file, int line) {
    fgk_set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return -1;
}

extern "C" int fgk_init(void) {
    if (g_inited) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        fgk_set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices"
                                                                      : hipGetErrorString(e));
        return -1;
    }
    FGK_HIP(hipSetDevice(0));
    g_inited = 1;
    return 0;
}

extern "C" int fgk_malloc(void **ptr, size_t bytes) {
    *ptr = NULL;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = NULL;
        fgk_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return -1;
    }
    return 0;
}

extern "C" int fgk_free(void *ptr) {
    if (ptr) FGK_HIP(hipFree(ptr));
    return 0;
}

extern "C" int fgk_host_alloc(void **ptr, size_t bytes) {
    *ptr = NULL;
    if (bytes == 0) bytes = 256;
    FGK_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

extern "C" int fgk_host_free(void *ptr) {
    if (ptr) FGK_HIP(hipHostFree(ptr));
    return 0;
}

extern "C" int fgk_h2d(void *dst, const void *src, size_t bytes, fgk_stream_t s) {
    if (!bytes) return 0;
    FGK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
    return 0;
}

extern "C" int fgk_d2h(void *dst, const void *src, size_t bytes, fgk_stream_t s) {
    if (!bytes) return 0;
    FGK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    return 0;
}

extern "C" int fgk_d2d(void *dst, const void *src, size_t bytes, fgk_stream_t s) {
    if (!bytes) return 0;
    FGK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
    return 0;
}

// zero fill with 16-byte stores; a kernel rather than a memset node so that
// captured graphs replay it correctly
__global__ void k_fill_zero16(uint4 *p, size_t n16) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_fill_zero1(unsigned char *p, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" int fgk_fill_zero(void *dst, size_t bytes, fgk_stream_t s) {
    if (!bytes) return 0;
    uintptr_t a = (uintptr_t)dst;
    size_t head = (16 - (a & 15)) & 15;
    if (head > bytes) head = bytes;
    if (head) {
        hipLaunchKernelGGL(k_fill_zero1, dim3(1), dim3(64), 0, (hipStream_t)s,
                           (unsigned char *)dst, head);
    }
    size_t rest = bytes - head;
    size_t n16 = rest / 16;
    if (n16) {
        size_t blocks = (n16 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_fill_zero16, dim3((unsigned)blocks), dim3(256), 0,
                           (hipStream_t)s, (uint4 *)((char *)dst + head), n16);
    }
    size_t tail = rest - n16 * 16;
    if (tail) {
        hipLaunchKernelGGL(k_fill_zero1, dim3(1), dim3(64), 0, (hipStream_t)s,
                           (unsigned char *)dst + head + n16 * 16, tail);
    }
    FGK_HIP(hipGetLastError());
    return 0;
}

extern "C" int fgk_stream_create(fgk_stream_t *s) {
    hipStream_t h;
    FGK_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = (fgk_stream_t)h;
    return 0;
}

extern "C" int fgk_stream_destroy(fgk_stream_t s) {
    if (s) FGK_HIP(hipStreamDestroy((hipStream_t)s));
    return 0;
}

extern "C" int fgk_stream_sync(fgk_stream_t s) {
    FGK_HIP(hipStreamSynchronize((hipStream_t)s));
    return 0;
}

extern "C" int fgk_event_create(fgk_event_t *e) {
    hipEvent_t h;
    FGK_HIP(hipEventCreate(&h));
    *e = (fgk_event_t)h;
    return 0;
}

// an event that only orders work (no timestamps)
extern "C" int fgk_event_create_dep(fgk_event_t *e) {
    hipEvent_t h;
    FGK_HIP(hipEventCreateWithFlags(&h, hipEventDisableTiming));
    *e = (fgk_event_t)h;
    return 0;
}

extern "C" int fgk_stream_wait_event(fgk_stream_t s, fgk_event_t e) {
    FGK_HIP(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0));
    return 0;
}

extern "C" int fgk_event_destroy(fgk_event_t e) {
    if (e) FGK_HIP(hipEventDestroy((hipEvent_t)e));
    return 0;
}

extern "C" int fgk_event_record(fgk_event_t e, fgk_stream_t s) {
    FGK_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return 0;
}

extern "C" size_t fgk_mem_free(void) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return fr;
}

extern "C" int fgk_event_sync(fgk_event_t e) {
    FGK_HIP(hipEventSynchronize((hipEvent_t)e));
    return 0;
}

extern "C" int fgk_event_elapsed_ms(float *ms, fgk_event_t a, fgk_event_t b) {
    FGK_HIP(hipEventSynchronize((hipEvent_t)b));
    FGK_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return 0;
}

extern "C" int fgk_device_info(char *name, size_t len, int *cu_count, size_t *mem_bytes) {
    hipDeviceProp_t prop;
    FGK_HIP(hipGetDeviceProperties(&prop, 0));
    if (name && len) {
        snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (mem_bytes) *mem_bytes = prop.totalGlobalMem;
    return 0;
}

extern "C" int fgk_graph_begin(fgk_stream_t s) {
    FGK_HIP(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal));
    return 0;
}

extern "C" int fgk_graph_end(fgk_stream_t s, fgk_graph_t *g) {
    hipGraph_t graph;
    FGK_HIP(hipStreamEndCapture((hipStream_t)s, &graph));
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, NULL, NULL, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fgk_hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    *g = (fgk_graph_t)exec;
    return 0;
}

extern "C" int fgk_graph_launch(fgk_graph_t g, fgk_stream_t s) {
    FGK_HIP(hipGraphLaunch((hipGraphExec_t)g, (hipStream_t)s));
    return 0;
}

extern "C" int fgk_graph_destroy(fgk_graph_t g) {
    if (g) FGK_HIP(hipGraphExecDestroy((hipGraphExec_t)g));
    return 0;
}

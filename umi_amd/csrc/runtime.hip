// Runtime part of the C ABI: device memory, copies, streams, event timers,
// graph capture and the raw launch used by the generic Execute<> engine.
// Replaces the CUDA runtime calls of src/SyncedMemory.h:19,22,65,74 and
// src/UMI_gpu.h:376-380,569,610 (johnjohnlin/UMI).
#include "common.h"
#include <string.h>

#include <algorithm>

namespace umi {

static thread_local char g_last_error[512] = "no error";

umi_status set_error(umi_status code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace umi

using namespace umi;

namespace {
// Plain streaming copy, 16 B per lane; used as the HBM byte-count reference
// for the profiler's FETCH_SIZE / WRITE_SIZE counters.
__global__ void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
// Stream-ordered fill of 32-bit words (a kernel, not hipMemsetAsync, so it is
// ordered with the caller's other work and safe inside captured graphs).
__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
}  // namespace

struct umi_timer_s {
    hipEvent_t start, stop;
};
struct umi_graph_s {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

const char *umi_last_error(void) { return g_last_error; }

umi_status umi_device_count(int *n) {
    if (!n) return set_error(UMI_ERR_INVALID, "umi_device_count: null pointer");
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        return set_error(UMI_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return UMI_OK;
}

umi_status umi_set_device(int device) { return hip_check(hipSetDevice(device), "hipSetDevice"); }

umi_status umi_device_name(char *buf, size_t len) {
    if (!buf || len == 0) return set_error(UMI_ERR_INVALID, "umi_device_name: empty buffer");
    int dev = 0;
    UMI_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    hipDeviceProp_t p;
    UMI_TRY(hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties"));
    // some drivers report an empty marketing name: fall back to the architecture and the board's size
    snprintf(buf, len, "%s (%s, %d CUs, %.0f GiB, %d MHz)", p.name[0] ? p.name : "AMD GPU", p.gcnArchName,
             p.multiProcessorCount, (double)p.totalGlobalMem / (1u << 30), p.clockRate / 1000);
    return UMI_OK;
}

umi_status umi_malloc(void **ptr, size_t bytes) {
    if (!ptr) return set_error(UMI_ERR_INVALID, "umi_malloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return UMI_OK;
    return hip_check(hipMalloc(ptr, bytes), "hipMalloc");
}

umi_status umi_free(void *ptr) {
    if (!ptr) return UMI_OK;
    return hip_check(hipFree(ptr), "hipFree");
}

umi_status umi_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return set_error(UMI_ERR_INVALID, "umi_host_alloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return UMI_OK;
    return hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}

umi_status umi_host_free(void *ptr) {
    if (!ptr) return UMI_OK;
    return hip_check(hipHostFree(ptr), "hipHostFree");
}

umi_status umi_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    if (bytes == 0) return UMI_OK;
    if (!dst || !src) return set_error(UMI_ERR_INVALID, "umi_memcpy: null pointer");
    if (kind < 0 || kind > 4) return set_error(UMI_ERR_INVALID, "umi_memcpy: bad kind %d", kind);
    return hip_check(hipMemcpy(dst, src, bytes, (hipMemcpyKind)kind), "hipMemcpy");
}

umi_status umi_memcpy_async(void *dst, const void *src, size_t bytes, int kind, umi_stream s) {
    if (bytes == 0) return UMI_OK;
    if (!dst || !src) return set_error(UMI_ERR_INVALID, "umi_memcpy_async: null pointer");
    if (kind < 0 || kind > 4) return set_error(UMI_ERR_INVALID, "umi_memcpy_async: bad kind %d", kind);
    return hip_check(hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, as_stream(s)),
                     "hipMemcpyAsync");
}

umi_status umi_memset(void *ptr, int value, size_t bytes) {
    if (bytes == 0) return UMI_OK;
    if (!ptr) return set_error(UMI_ERR_INVALID, "umi_memset: null pointer");
    return hip_check(hipMemset(ptr, value, bytes), "hipMemset");
}

umi_status umi_fill_u32(void *ptr, uint32_t value, size_t count, umi_stream s) {
    if (count == 0) return UMI_OK;
    if (!ptr || ((uintptr_t)ptr & 3)) return set_error(UMI_ERR_INVALID, "umi_fill_u32: null or unaligned pointer");
    const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, as_stream(s), (uint32_t *)ptr, value, count);
    return launch_check("fill_u32_kernel");
}

umi_status umi_copy_d2d(void *dst, const void *src, size_t bytes, umi_stream s) {
    if (bytes == 0) return UMI_OK;
    if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return set_error(UMI_ERR_INVALID, "umi_copy_d2d: pointers and size must be 16-byte multiples");
    const size_t n16 = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, as_stream(s), (const uint4 *)src, (uint4 *)dst, n16);
    return launch_check("copy16_kernel");
}

umi_status umi_synchronize(void) { return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

umi_status umi_stream_create(umi_stream *s) {
    if (!s) return set_error(UMI_ERR_INVALID, "umi_stream_create: null pointer");
    hipStream_t st;
    UMI_TRY(hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate"));
    *s = reinterpret_cast<umi_stream>(st);
    return UMI_OK;
}

umi_status umi_stream_destroy(umi_stream s) {
    if (!s) return UMI_OK;
    return hip_check(hipStreamDestroy(as_stream(s)), "hipStreamDestroy");
}

umi_status umi_stream_synchronize(umi_stream s) {
    return hip_check(hipStreamSynchronize(as_stream(s)), "hipStreamSynchronize");
}

umi_status umi_stream_wait(umi_stream waiter, umi_stream signaller) {
    hipEvent_t e;
    UMI_TRY(hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"));
    umi_status rc = hip_check(hipEventRecord(e, as_stream(signaller)), "hipEventRecord");
    if (rc == UMI_OK) rc = hip_check(hipStreamWaitEvent(as_stream(waiter), e, 0), "hipStreamWaitEvent");
    const umi_status rd = hip_check(hipEventDestroy(e), "hipEventDestroy");  // the enqueued wait keeps its event
    return rc != UMI_OK ? rc : rd;
}

umi_status umi_launch_kernel(const void *func, unsigned gx, unsigned gy, unsigned gz, unsigned bx,
                             unsigned by, unsigned bz, size_t shmem, void **args, umi_stream s) {
    if (!func) return set_error(UMI_ERR_INVALID, "umi_launch_kernel: null kernel");
    if ((size_t)gx * gy * gz == 0) return UMI_OK;
    if ((size_t)bx * by * bz == 0 || (size_t)bx * by * bz > 1024)
        return set_error(UMI_ERR_INVALID, "umi_launch_kernel: block of %u x %u x %u threads", bx, by, bz);
    if (shmem > 160 * 1024)
        return set_error(UMI_ERR_UNSUPPORTED, "umi_launch_kernel: %zu B of LDS exceeds 160 KiB", shmem);
    return hip_check(hipLaunchKernel(func, dim3(gx, gy, gz), dim3(bx, by, bz), args, shmem, as_stream(s)),
                     "hipLaunchKernel");
}

umi_status umi_timer_create(umi_timer *t) {
    if (!t) return set_error(UMI_ERR_INVALID, "umi_timer_create: null pointer");
    umi_timer_s *p = new umi_timer_s;
    if (hipEventCreate(&p->start) != hipSuccess || hipEventCreate(&p->stop) != hipSuccess) {
        delete p;
        return set_error(UMI_ERR_HIP, "hipEventCreate failed");
    }
    *t = p;
    return UMI_OK;
}

umi_status umi_timer_destroy(umi_timer t) {
    if (!t) return UMI_OK;
    hipEventDestroy(t->start);
    hipEventDestroy(t->stop);
    delete t;
    return UMI_OK;
}

umi_status umi_timer_start(umi_timer t, umi_stream s) {
    if (!t) return set_error(UMI_ERR_INVALID, "umi_timer_start: null timer");
    return hip_check(hipEventRecord(t->start, as_stream(s)), "hipEventRecord");
}

umi_status umi_timer_stop(umi_timer t, umi_stream s) {
    if (!t) return set_error(UMI_ERR_INVALID, "umi_timer_stop: null timer");
    return hip_check(hipEventRecord(t->stop, as_stream(s)), "hipEventRecord");
}

umi_status umi_timer_elapsed_ms(umi_timer t, float *ms) {
    if (!t || !ms) return set_error(UMI_ERR_INVALID, "umi_timer_elapsed_ms: null pointer");
    UMI_TRY(hip_check(hipEventSynchronize(t->stop), "hipEventSynchronize"));
    return hip_check(hipEventElapsedTime(ms, t->start, t->stop), "hipEventElapsedTime");
}

umi_status umi_graph_begin(umi_stream s) {
    if (!s) return set_error(UMI_ERR_INVALID, "umi_graph_begin: capture needs a created stream");
    return hip_check(hipStreamBeginCapture(as_stream(s), hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

umi_status umi_graph_end(umi_stream s, umi_graph *g) {
    if (!s || !g) return set_error(UMI_ERR_INVALID, "umi_graph_end: null argument");
    umi_graph_s *p = new umi_graph_s;
    hipError_t e = hipStreamEndCapture(as_stream(s), &p->graph);
    if (e != hipSuccess) {
        delete p;
        return hip_check(e, "hipStreamEndCapture");
    }
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(p->graph);
        delete p;
        return hip_check(e, "hipGraphInstantiate");
    }
    *g = p;
    return UMI_OK;
}

umi_status umi_graph_launch(umi_graph g, umi_stream s) {
    if (!g) return set_error(UMI_ERR_INVALID, "umi_graph_launch: null graph");
    return hip_check(hipGraphLaunch(g->exec, as_stream(s)), "hipGraphLaunch");
}

umi_status umi_graph_destroy(umi_graph g) {
    if (!g) return UMI_OK;
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    return UMI_OK;
}

}  // extern "C"

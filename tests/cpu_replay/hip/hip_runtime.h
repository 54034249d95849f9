// Host emulation of the parts of the HIP runtime and of the device language that libh2g's host
// code and its generic (non-MFMA) kernels use -- TEST INFRASTRUCTURE ONLY.
//
// The CPU replay compiles the library's own api.hip, kernels.hip and method.hip with g++ against
// this header (it shadows <hip/hip_runtime.h>) under AddressSanitizer, so that every global and
// LDS access of those kernels is checked against its allocation:
//   * device memory is malloc'ed with the exact size asked for (ASan redzones on both sides);
//   * dynamic LDS is a fixed array whose bytes beyond the launch's dynamic size are poisoned;
//   * static __shared__ arrays become function-local statics (ASan-instrumented globals);
//   * a workgroup of B threads runs as B OS threads; __syncthreads() is a std::barrier, so a
//     barrier that not every thread reaches hangs (and the watchdog of the replay reports it);
//   * workgroups run one after another; streams, events and graphs are executed in order
//     (a captured graph records its launches and copies and replays them on hipGraphLaunch).
#pragma once

#include <sanitizer/asan_interface.h>

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

// ------------------------------------------------------------------------------------------
// device language
// ------------------------------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

inline thread_local dim3 threadIdx, blockIdx;
inline dim3 blockDim, gridDim;

namespace h2g_cpu {
inline std::barrier<> *g_bar = nullptr;
constexpr size_t LDS_BYTES = 160 * 1024;
}  // namespace h2g_cpu

inline void __syncthreads() { h2g_cpu::g_bar->arrive_and_wait(); }

template <class A, class B>
inline std::common_type_t<A, B> min(A a, B b) { return a < b ? a : b; }
template <class A, class B>
inline std::common_type_t<A, B> max(A a, B b) { return a < b ? b : a; }
inline void sincos(double x, double *s, double *c) { *s = std::sin(x); *c = std::cos(x); }
// device atomics on host memory: the workgroup's threads run concurrently
inline int atomicOr(int *p, int v) { return std::atomic_ref<int>(*p).fetch_or(v); }
inline void sincospi(double x, double *s, double *c) { *s = std::sin(M_PI * x); *c = std::cos(M_PI * x); }

// the dynamic LDS arrays the kernels declare (extern __shared__ unsigned char smem_raw[]),
// in namespace h2g and in its unnamed namespace of each translation unit
namespace h2g {
alignas(256) inline unsigned char smem_raw[h2g_cpu::LDS_BYTES];
inline unsigned char *smem_raw_named() { return smem_raw; }
namespace {
alignas(256) unsigned char smem_raw[h2g_cpu::LDS_BYTES];
inline unsigned char *smem_raw_anon() { return smem_raw; }
}  // namespace
}  // namespace h2g

// ------------------------------------------------------------------------------------------
// runtime
// ------------------------------------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorMemoryAllocation = 2, hipErrorNotReady = 600 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2

namespace h2g_cpu {
using Op = std::function<void()>;
struct Stream {
    std::vector<Op> *capture = nullptr;  // recording into a graph
};
struct Graph {
    std::vector<Op> ops;
};
inline Stream g_null_stream;
inline Stream *stream_of(void *s) { return s ? static_cast<Stream *>(s) : &g_null_stream; }
inline void submit(void *s, Op op) {
    Stream *st = stream_of(s);
    if (st->capture)
        st->capture->push_back(std::move(op));
    else
        op();
}
inline std::atomic<long> g_launches{0};
}  // namespace h2g_cpu

typedef struct h2g_cpu::Stream *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
typedef struct h2g_cpu::Graph *hipGraph_t;
typedef struct h2g_cpu::Graph *hipGraphExec_t;

inline const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "hipSuccess" : "hip error (host replay)"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) {
    *p = std::malloc(n);
    return *p ? hipSuccess : hipErrorMemoryAllocation;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
    if (n) std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    if (n) std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st = nullptr) {
    h2g_cpu::submit(st, [=] { if (n) std::memcpy(d, s, n); });
    return hipSuccess;
}
inline hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind,
                                   hipStream_t st = nullptr) {
    h2g_cpu::submit(st, [=] {
        for (size_t r = 0; r < h; ++r) std::memcpy((char *)d + r * dp, (const char *)s + r * sp, w);
    });
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = new h2g_cpu::Stream; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = reinterpret_cast<hipEvent_t>(new char); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete reinterpret_cast<char *>(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    auto *g = new h2g_cpu::Graph;
    h2g_cpu::stream_of(s)->capture = &g->ops;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    auto *st = h2g_cpu::stream_of(s);
    // the Graph owns the vector the stream recorded into
    *g = reinterpret_cast<hipGraph_t>(reinterpret_cast<char *>(st->capture) - offsetof(h2g_cpu::Graph, ops));
    st->capture = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t) {
    *e = new h2g_cpu::Graph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    for (auto &op : e->ops) h2g_cpu::submit(s, op);
    return hipSuccess;
}

// ------------------------------------------------------------------------------------------
// kernel launch
// ------------------------------------------------------------------------------------------
namespace h2g_cpu {
// runs the grid: one OS thread per work-item, workgroups one after another
inline void run_grid(dim3 g, dim3 b, size_t dyn_lds, unsigned char *lds_a, unsigned char *lds_b,
                     const std::function<void()> &body) {
    if (dyn_lds > LDS_BYTES) {
        std::fprintf(stderr, "host replay: dynamic LDS of %zu bytes exceeds %zu\n", dyn_lds, LDS_BYTES);
        std::abort();
    }
    const unsigned nt = b.x * b.y * b.z;
    if (nt == 0 || nt > 1024) {
        std::fprintf(stderr, "host replay: workgroup of %u threads\n", nt);
        std::abort();
    }
    for (unsigned char *lds : {lds_a, lds_b}) {
        ASAN_UNPOISON_MEMORY_REGION(lds, LDS_BYTES);
        std::memset(lds, 0xA5, LDS_BYTES);  // LDS is not zeroed between workgroups
        ASAN_POISON_MEMORY_REGION(lds + dyn_lds, LDS_BYTES - dyn_lds);
    }
    std::barrier<> bar((std::ptrdiff_t)nt);
    g_bar = &bar;
    blockDim = b;
    gridDim = g;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            threadIdx = dim3(t % b.x, (t / b.x) % b.y, t / (b.x * b.y));
            for (unsigned bz = 0; bz < g.z; ++bz)
                for (unsigned by = 0; by < g.y; ++by)
                    for (unsigned bx = 0; bx < g.x; ++bx) {
                        blockIdx = dim3(bx, by, bz);
                        body();
                        bar.arrive_and_wait();  // the next workgroup reuses the LDS
                    }
        });
    for (auto &x : th) x.join();
    g_bar = nullptr;
    for (unsigned char *lds : {lds_a, lds_b}) ASAN_UNPOISON_MEMORY_REGION(lds, LDS_BYTES);
    g_launches++;
}
}  // namespace h2g_cpu

namespace {
template <class K, class... A>
void h2g_cpu_launch(hipStream_t s, dim3 g, dim3 b, size_t dyn_lds, K kern, A... args) {
    h2g_cpu::submit(s, [=] {
        h2g_cpu::run_grid(g, b, dyn_lds, h2g::smem_raw_named(), h2g::smem_raw_anon(), [&] { kern(args...); });
    });
}
}  // namespace

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
    h2g_cpu_launch((hipStream_t)(stream), dim3(grid), dim3(block), (size_t)(lds), kernel, ##__VA_ARGS__)

// CPU execution of the HIP subset used by qcxms_amd/csrc, for bounds checking
// the kernels on the host (built with clang++ -fsanitize=address).
//
// Every GPU thread of a block is an OS thread; blocks of a launch run one
// after the other.  __syncthreads is a block barrier, __shfl_xor, __shfl and the
// FP64 MFMA 16x16x4 tile product exchange operands through per-wavefront
// slots between two wavefront barriers.  Threads that leave a kernel early
// drop out of the barriers, as exited wavefronts do on the GPU.  Dynamic LDS
// is a heap block of exactly the launch's size, so an access beyond it is
// reported by the sanitizer; it and the static LDS variables start every
// block filled with 0xff bytes, so a read before a write shows up as NaN or -1.  Device memory is host memory; copies and fills run
// synchronously; stream capture records the calls and a graph launch replays
// them.  This is test infrastructure only: it is never linked into the
// library that the Python package and the C ABI load.
#pragma once

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __constant__
// Static LDS variables live in one section so that every block can start
// with it filled with 0xff bytes (NaN doubles, -1 integers), as the GPU
// starts a workgroup with whatever the last one left in LDS.
#define __shared__ static __attribute__((section("qx_emu_lds")))
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(X) ((const void*)&(X))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct double2 {
    double x, y;
};
struct int2 {
    int x, y;
};
inline double2 make_double2(double x, double y) { return {x, y}; }
inline int2 make_int2(int x, int y) { return {x, y}; }

inline thread_local dim3 threadIdx;
inline thread_local dim3 blockIdx;
inline dim3 blockDim;
inline dim3 gridDim;

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorEmu = 1 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
struct qx_emu_stream {};
typedef qx_emu_stream* hipStream_t;
struct qx_emu_event {
    std::chrono::steady_clock::time_point t;
};
typedef qx_emu_event* hipEvent_t;
struct qx_emu_graph {
    std::vector<std::function<void()>> ops;
};
typedef qx_emu_graph* hipGraph_t;
typedef qx_emu_graph* hipGraphExec_t;

namespace qx_emu {

constexpr int kMaxThreads = 1024;
constexpr size_t kMaxLds = 160 * 1024;
constexpr int kTimeoutSeconds = 300;

// Barrier whose participant count can shrink while it is in use.  A block
// barrier also checks that every thread of one barrier generation arrives
// from the same source line: a __syncthreads reached from two places in one
// round means control flow around a barrier that is not uniform across the
// block, which the GPU does not report (its barrier counts arrivals from
// anywhere) and which can desynchronise the waves or hang them.
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, arrived = 0;
    unsigned long gen = 0;
    const char* site_file = nullptr;
    int site_line = 0;
    unsigned site_thread = 0;
    void reset(int n)
    {
        std::lock_guard<std::mutex> l(m);
        expected = n;
        arrived = 0;
    }
    void wait(const char* what, const char* file = nullptr, int line = 0)
    {
        std::unique_lock<std::mutex> l(m);
        const unsigned long g = gen;
        if (file) {
            if (arrived == 0) {
                site_file = file;
                site_line = line;
                site_thread = threadIdx.x;
            } else if (line != site_line || std::strcmp(file, site_file) != 0) {
                std::fprintf(stderr,
                             "cpu emulation: divergent %s barrier (block %u): thread %u at %s:%d, "
                             "thread %u at %s:%d\n",
                             what, blockIdx.x, site_thread, site_file, site_line, threadIdx.x, file, line);
                std::abort();
            }
        }
        if (++arrived == expected) {
            arrived = 0;
            ++gen;
            cv.notify_all();
            return;
        }
        if (!cv.wait_for(l, std::chrono::seconds(kTimeoutSeconds), [&] { return gen != g; })) {
            std::fprintf(stderr, "cpu emulation: %s barrier not reached by all threads (block %u thread %u, %d of %d)\n",
                         what, blockIdx.x, threadIdx.x, arrived, expected);
            std::abort();
        }
    }
    void drop()
    {
        std::lock_guard<std::mutex> l(m);
        --expected;
        if (arrived > 0 && arrived == expected) {
            arrived = 0;
            ++gen;
            cv.notify_all();
        }
    }
};

struct Wave {
    Barrier bar;
    double a[64], b[64];
};

inline Barrier g_block, g_end;
inline Wave g_wave[kMaxThreads / 64];
inline unsigned char* g_dyn = nullptr;
inline bool g_capturing = false;
inline std::vector<std::function<void()>> g_capture;

inline Wave& wave() { return g_wave[threadIdx.x / 64]; }

template <class T>
inline T shfl_xor(T v, int mask, int width)
{
    static_assert(sizeof(T) <= sizeof(double), "shuffle of a wider type");
    const int l = threadIdx.x & 63;
    Wave& w = wave();
    double slot = 0.0;
    std::memcpy(&slot, &v, sizeof(T));
    w.a[l] = slot;
    w.bar.wait("wavefront shuffle");
    const int src = (l & ~(width - 1)) | ((l ^ mask) & (width - 1));
    const double got = w.a[src];
    w.bar.wait("wavefront shuffle");
    T out;
    std::memcpy(&out, &got, sizeof(T));
    return out;
}

// __shfl: every lane receives the value of lane src (within width lanes)
template <class T>
inline T shfl(T v, int src, int width)
{
    static_assert(sizeof(T) <= sizeof(double), "shuffle of a wider type");
    const int l = threadIdx.x & 63;
    Wave& w = wave();
    double slot = 0.0;
    std::memcpy(&slot, &v, sizeof(T));
    w.a[l] = slot;
    w.bar.wait("wavefront shuffle");
    const int from = (l & ~(width - 1)) | (src & (width - 1));
    const double got = w.a[from];
    w.bar.wait("wavefront shuffle");
    T out;
    std::memcpy(&out, &got, sizeof(T));
    return out;
}

// v_readlane: every lane of the wave receives lane l's value; l must be the
// same in all lanes of the wave.
inline int readlane(int v, int l)
{
    const int me = threadIdx.x & 63;
    Wave& w = wave();
    w.a[me] = (double)v;
    w.b[me] = (double)l;
    w.bar.wait("lane read");
    if (l < 0 || l > 63 || w.b[0] != (double)l) {
        std::fprintf(stderr, "cpu emulation: lane read of lane %d not uniform (thread %u)\n", l, threadIdx.x);
        std::abort();
    }
    const int out = (int)w.a[l];
    w.bar.wait("lane read");
    return out;
}

inline long long double_as_longlong(double d)
{
    long long b;
    std::memcpy(&b, &d, sizeof b);
    return b;
}

inline double longlong_as_double(long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

typedef double d4 __attribute__((ext_vector_type(4)));

// D[i][j] = C + sum_k A[i][k] B[k][j] over a 16x16x4 tile: lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[4 r + l / 16][l % 16], r = 0..3.
inline d4 mfma_f64_16x16x4(double a, double b, d4 c, int, int, int)
{
    const int l = threadIdx.x & 63;
    Wave& w = wave();
    w.a[l] = a;
    w.b[l] = b;
    w.bar.wait("MFMA operand exchange");
    d4 d = c;
    const int j = l % 16;
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * r + l / 16;
        double s = d[r];
        for (int k = 0; k < 4; ++k) s += w.a[i + 16 * k] * w.b[16 * k + j];
        d[r] = s;
    }
    w.bar.wait("MFMA operand exchange");
    return d;
}

inline void* dyn_lds() { return g_dyn; }

extern "C" char __start_qx_emu_lds[], __stop_qx_emu_lds[];
inline void poison_static_lds() { std::memset(__start_qx_emu_lds, 0xff, __stop_qx_emu_lds - __start_qx_emu_lds); }

template <class... P, class... A>
void launch(void (*k)(P...), dim3 grid, dim3 block, size_t shmem, A... args)
{
    const int nt = (int)(block.x * block.y * block.z);
    if (nt <= 0 || nt > kMaxThreads || block.y != 1 || block.z != 1 || grid.y != 1 || grid.z != 1) {
        std::fprintf(stderr, "cpu emulation: unsupported launch shape\n");
        std::abort();
    }
    if (shmem > kMaxLds) {
        std::fprintf(stderr, "cpu emulation: %zu bytes of dynamic LDS exceed 160 KiB\n", shmem);
        std::abort();
    }
    gridDim = grid;
    blockDim = block;
    g_end.reset(nt);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            threadIdx = dim3(t);
            for (unsigned b = 0; b < grid.x; ++b) {
                blockIdx = dim3(b);
                if (t == 0) {
                    g_block.reset(nt);
                    for (int wv = 0; wv < (nt + 63) / 64; ++wv)
                        g_wave[wv].bar.reset(std::min(64, nt - 64 * wv));
                    std::free(g_dyn);
                    g_dyn = (unsigned char*)std::malloc(shmem ? shmem : 1);
                    std::memset(g_dyn, 0xff, shmem ? shmem : 1);
                    poison_static_lds();
                }
                g_end.wait("block start");
                k(args...);
                g_block.drop();
                wave().bar.drop();
                g_end.wait("block end");
            }
        });
    for (auto& x : th) x.join();
    std::free(g_dyn);
    g_dyn = nullptr;
}

inline void run_or_record(std::function<void()> f)
{
    if (g_capturing)
        g_capture.push_back(std::move(f));
    else
        f();
}

}  // namespace qx_emu

#define __syncthreads() qx_emu::g_block.wait("block", __FILE__, __LINE__)
#define __shfl_xor(v, m, w) qx_emu::shfl_xor((v), (m), (w))
#define __shfl(v, l, w) qx_emu::shfl((v), (l), (w))
#define __builtin_amdgcn_mfma_f64_16x16x4f64 qx_emu::mfma_f64_16x16x4
#define __ffsll(x) __builtin_ffsll(x)
#define __builtin_amdgcn_readlane(v, l) qx_emu::readlane((v), (l))
#define __double_as_longlong(d) qx_emu::double_as_longlong(d)
#define __longlong_as_double(b) qx_emu::longlong_as_double(b)

template <class... P, class... A>
inline void hipLaunchKernelGGL(void (*k)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t, A... args)
{
    qx_emu::run_or_record([=] { qx_emu::launch(k, grid, block, shmem, args...); });
}

inline hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "error in the CPU emulation"; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int v)
{
    return v <= (int)qx_emu::kMaxLds ? hipSuccess : hipErrorEmu;
}
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 1 };
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int)
{
    *v = (int)qx_emu::kMaxLds;
    return hipSuccess;
}
// Static LDS per kernel: the emulation keeps every kernel's __shared__
// objects in one section, so it reports the MI355X build's figure for the
// largest kernels (46,480 bytes for the SCC and persistent MD kernels),
// rounded up.
struct hipFuncAttributes {
    size_t sharedSizeBytes;
};
inline hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*)
{
    a->sharedSizeBytes = 47 * 1024;
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = new qx_emu_stream;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n)
{
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorEmu;
}
inline hipError_t hipFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    if (n) std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t)
{
    qx_emu::run_or_record([=] {
        if (n) std::memcpy(d, s, n);
    });
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t)
{
    qx_emu::run_or_record([=] {
        if (n) std::memset(d, v, n);
    });
    return hipSuccess;
}
inline hipError_t hipMemcpyToSymbol(const void* sym, const void* src, size_t n, size_t off = 0,
                                    hipMemcpyKind = hipMemcpyHostToDevice)
{
    std::memcpy((char*)sym + off, src, n);
    return hipSuccess;
}
// the address of a device symbol is its host address here
inline hipError_t hipGetSymbolAddress(void** p, const void* sym)
{
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
inline hipError_t hipMemcpyFromSymbol(void* dst, const void* sym, size_t n, size_t off = 0,
                                      hipMemcpyKind = hipMemcpyDeviceToHost)
{
    std::memcpy(dst, (const char*)sym + off, n);
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = new qx_emu_event;
    return hipSuccess;
}
#define hipEventDisableTiming 0x2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// every emulated launch completes in program order, so a cross-stream wait
// has nothing to wait for
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    qx_emu::run_or_record([=] { e->t = std::chrono::steady_clock::now(); });
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    qx_emu::g_capturing = true;
    qx_emu::g_capture.clear();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    qx_emu::g_capturing = false;
    *g = new qx_emu_graph{std::move(qx_emu::g_capture)};
    qx_emu::g_capture.clear();
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t)
{
    *e = new qx_emu_graph{g->ops};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    for (auto& f : e->ops) f();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete e;
    return hipSuccess;
}

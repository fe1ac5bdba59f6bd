// CPU simulation of the subset of the HIP runtime and device language that
// the library uses (test infrastructure only; never part of the product).
//
// The library's kernel source is compiled for the host with this header in
// place of the ROCm one.  A launch runs every workgroup in turn; the threads
// of a workgroup are fibers on one OS thread, switched at __syncthreads()
// (all live threads of the block) and at wave collectives (__shfl,
// __shfl_xor, __ballot, the matrix-core product: all live lanes of the
// 64-lane wave).  Device and pinned host allocations are fenced by
// inaccessible guard pages, so an out-of-bounds access traps at once; the
// trap handler names the kernel, workgroup, thread and nearest allocation.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __constant__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_DYNAMIC_SHARED(type, var) type *var = reinterpret_cast<type *>(sim::dynamic_smem());
#define HIP_SYMBOL(x) (&(x))

// ---- vector types ----
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct __attribute__((aligned(8))) float2 {
    float x, y;
};
struct __attribute__((aligned(16))) float4 {
    float x, y, z, w;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }

inline float __uint_as_float(unsigned u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline unsigned __float_as_uint(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
// device min/max of mixed arithmetic types (the usual conversions)
template <class A, class B>
inline auto min(A a, B b) -> decltype(a + b)
{
    return b < a ? b : a;
}
template <class A, class B>
inline auto max(A a, B b) -> decltype(a + b)
{
    return a < b ? b : a;
}
// the device library's FP64 sin/cos of pi * x
inline void sincospi(double x, double *s, double *c)
{
    *s = std::sin(M_PI * x);
    *c = std::cos(M_PI * x);
}

// ---- runtime types ----
typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorNotReady = 600,
    hipErrorOutOfMemory = 2,
    hipErrorMemoryAllocation = 2,
    hipErrorInvalidDevice = 101,
    hipErrorNoDevice = 100,
    hipErrorNotSupported = 801,
    hipErrorUnknown = 999,
};
struct ihipStream_t;
struct sim_event;
struct sim_graph;
typedef ihipStream_t *hipStream_t;
typedef sim_event *hipEvent_t;
typedef sim_graph *hipGraph_t;
typedef sim_graph *hipGraphExec_t;
typedef int hipDevice_t;
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
enum { hipStreamDefault = 0, hipStreamNonBlocking = 1 };
enum { hipHostMallocDefault = 0 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };

// ---- runtime API (sim_runtime.cpp) ----
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipStreamGetDevice(hipStream_t s, hipDevice_t *d);
// one device, one context
struct ihipCtx_t;
typedef struct ihipCtx_t *hipCtx_t;
hipError_t hipCtxGetCurrent(hipCtx_t *c);
hipError_t hipCtxPushCurrent(hipCtx_t c);
hipError_t hipCtxPopCurrent(hipCtx_t *c);
hipError_t hipCtxGetDevice(hipDevice_t *d);
hipError_t hipMalloc(void **p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc(reinterpret_cast<void **>(p), bytes); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags = 0);
template <class T>
inline hipError_t hipHostMalloc(T **p, size_t bytes, unsigned flags = 0)
{
    return hipHostMalloc(reinterpret_cast<void **>(p), bytes, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags = 0);
// a small simulated device: 4 compute units, 2 resident workgroups each
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 63 };
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t attr, int device);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *kernel, int threads, size_t smem);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);

// ---- device side ----
namespace sim {
struct Idx {
    unsigned x, y, z;
};
extern Idx g_thread, g_block, g_bdim, g_gdim;
void launch(const char *name, dim3 grid, dim3 block, size_t smem, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
            const std::function<void()> &body);
void syncthreads();
void *dynamic_smem();
// wave collectives: every live lane of the wave contributes `in`; returns
// the inputs of all 64 lanes (index by lane; lanes that have exited hold 0)
const void *wave_exchange(const void *in, size_t bytes);
// parks the lane until every live lane of its wave has arrived
void wave_barrier();
int lane_id();
}  // namespace sim

#define threadIdx (::sim::g_thread)
#define blockIdx (::sim::g_block)
#define blockDim (::sim::g_bdim)
#define gridDim (::sim::g_gdim)
#define warpSize 64

inline void __syncthreads() { sim::syncthreads(); }

template <class T>
inline T __shfl(T v, int src, int width = 64)
{
    const T *all = static_cast<const T *>(sim::wave_exchange(&v, sizeof(T)));
    const int lane = sim::lane_id();
    const int base = lane & ~(width - 1);
    return all[base + (((src % width) + width) % width)];
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
    const T *all = static_cast<const T *>(sim::wave_exchange(&v, sizeof(T)));
    const int lane = sim::lane_id();
    const int base = lane & ~(width - 1);
    return all[base + (((lane - base) ^ mask) & (width - 1))];
}
inline unsigned long long __ballot(int pred)
{
    const int p = pred ? 1 : 0;
    const int *all = static_cast<const int *>(sim::wave_exchange(&p, sizeof(int)));
    unsigned long long m = 0;
    for (int l = 0; l < 64; l++)
        if (all[l]) m |= 1ull << l;
    return m;
}
// single OS thread: read-modify-write is atomic between fibers
template <class T>
inline T atomicAdd(T *p, T v)
{
    const T old = *p;
    *p = old + v;
    return old;
}

#define __builtin_amdgcn_s_setprio(x) ((void)0)
#define __builtin_amdgcn_wave_barrier() ::sim::wave_barrier()
// used on wave-uniform values only
#define __builtin_amdgcn_readfirstlane(x) (x)
// fibers of one OS thread see each other's stores at once
#define __builtin_amdgcn_fence(order, scope) ((void)0)
// v_mfma_f32_16x16x32_bf16 on the wave (lane layout as the library documents it)
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, C, cbsz, abid, blgp) sim_mfma_16x16x32_bf16(A, B, C)
typedef __bf16 sim_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sim_f32x4 __attribute__((ext_vector_type(4)));
inline sim_f32x4 sim_mfma_16x16x32_bf16(sim_bf16x8 a, sim_bf16x8 b, sim_f32x4 c)
{
    struct Frag {
        sim_bf16x8 a, b;
    } mine{a, b};
    const Frag *all = static_cast<const Frag *>(sim::wave_exchange(&mine, sizeof(Frag)));
    const int l = sim::lane_id();
    sim_f32x4 d = c;
    for (int r = 0; r < 4; r++) {
        const int row = 4 * (l / 16) + r, col = l % 16;
        float acc = d[r];
        for (int k = 0; k < 32; k++) {
            // A: lane (row + 16 * (k / 8)) holds element k % 8; B: lane (col + 16 * (k / 8))
            const float av = (float)all[row + 16 * (k / 8)].a[k % 8];
            const float bv = (float)all[col + 16 * (k / 8)].b[k % 8];
            acc += av * bv;
        }
        d[r] = acc;
    }
    return d;
}

#define hipLaunchKernelGGL(K, G, BL, SH, S, ...) \
    (::sim::launch(#K, dim3(G), dim3(BL), (size_t)(SH), (S), nullptr, nullptr, [=]() { K(__VA_ARGS__); }))
#define hipExtLaunchKernelGGL(K, G, BL, SH, S, E0, E1, FL, ...) \
    (::sim::launch(#K, dim3(G), dim3(BL), (size_t)(SH), (S), (E0), (E1), [=]() { K(__VA_ARGS__); }))

// HIP on the CPU, for the host-side sanitizer harness of libkrt
// (tests/emu/build.py): the library's kernels and host code compiled by clang
// as ordinary C++ with AddressSanitizer and UBSan, so that every access a
// kernel makes is checked against the bounds of its host allocation.
//
// Execution model.  Every runtime call runs to completion when it is made, in
// the order the host makes it (streams and events only order; there is no
// overlap).  A kernel launch runs its workgroups one after another; the lanes
// of a workgroup are fibers on the calling thread, switched at every
// __syncthreads and wave operation (__ballot, __shfl*, MFMA): a rendezvous
// completes when every live lane of the workgroup (or wave) has reached it,
// and a lane that returns leaves the rendezvous sets, as a finished wavefront
// leaves s_barrier on the GPU.  __shared__ variables are function statics,
// shared by the lanes of the one workgroup that runs at a time.
//
// Wave operations must be reached by every live lane of the wave at the same
// point (uniform control flow), which is how the library uses them; when no
// lane can run and no rendezvous can complete, the process aborts with the
// kernel's name instead of hanging.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>

// ------------------------------------------------------------- qualifiers
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

using std::isnan;
using std::isinf;
using std::isfinite;
using std::signbit;

// ------------------------------------------------------------- vector types
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct alignas(16) float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct alignas(16) int4 { int x, y, z, w; };
struct uint2 { unsigned x, y; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
inline float2 make_float2(float x, float y) { return {x, y}; }
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }
inline uint2 make_uint2(unsigned x, unsigned y) { return {x, y}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return {x, y, z, w}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct krt_emu_idx {
    unsigned x, y, z;
};
extern krt_emu_idx threadIdx, blockIdx, blockDim, gridDim;

// integer min / max of the device library
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a < b ? b : a; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a < b ? b : a; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a < b ? b : a; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a < b ? b : a; }

// ------------------------------------------------------------- bit casts
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll((long long)v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }

// ------------------------------------------------------------- rendezvous
namespace krt_emu {
void syncthreads();
int syncthreads_count(int pred);
unsigned long long wave_ballot(int pred);
uint64_t wave_exchange(uint64_t mine, int src_lane);  // value of src_lane (0 if it has returned)
}  // namespace krt_emu

inline void __syncthreads() { krt_emu::syncthreads(); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline int __syncthreads_count(int pred) { return krt_emu::syncthreads_count(pred); }
inline unsigned long long __ballot(int pred) { return krt_emu::wave_ballot(pred); }
inline int __any(int pred) { return krt_emu::wave_ballot(pred) != 0ull; }
inline int __all(int pred) { return krt_emu::wave_ballot(!pred) == 0ull; }
// the wave's LDS accesses before it are visible to its lanes after it
inline void __builtin_amdgcn_wave_barrier() { (void)krt_emu::wave_ballot(0); }

template <class T>
inline T krt_emu_from_bits(uint64_t b) {
    static_assert(sizeof(T) <= 8, "wave exchange of at most 8 bytes");
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}
template <class T>
inline uint64_t krt_emu_to_bits(T v) {
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T __shfl(T v, int src, int width = 64) {
    const int lane = (int)(threadIdx.x & 63u);
    const int s = (lane & ~(width - 1)) + (src & (width - 1));
    return krt_emu_from_bits<T>(krt_emu::wave_exchange(krt_emu_to_bits(v), s));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    const int lane = (int)(threadIdx.x & 63u);
    int s = lane ^ mask;
    if ((s & ~(width - 1)) != (lane & ~(width - 1))) s = lane;
    return krt_emu_from_bits<T>(krt_emu::wave_exchange(krt_emu_to_bits(v), s));
}
template <class T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
    const int lane = (int)(threadIdx.x & 63u);
    int s = lane - (int)delta;
    if (s < (lane & ~(width - 1))) s = lane;
    return krt_emu_from_bits<T>(krt_emu::wave_exchange(krt_emu_to_bits(v), s));
}
template <class T>
inline T __shfl_down(T v, unsigned delta, int width = 64) {
    const int lane = (int)(threadIdx.x & 63u);
    int s = lane + (int)delta;
    if ((s & ~(width - 1)) != (lane & ~(width - 1))) s = lane;
    return krt_emu_from_bits<T>(krt_emu::wave_exchange(krt_emu_to_bits(v), s));
}

// v_mfma_f32_16x16x32_bf16 of one wave: D = A B + C, A 16 x 32, B 32 x 16.
// Lane l holds A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16] (j < 8)
// and D[4 (l / 16) + i][l % 16] (i < 4).
typedef __bf16 krt_emu_bf16x8 __attribute__((ext_vector_type(8)));
typedef float krt_emu_f32x4 __attribute__((ext_vector_type(4)));
namespace krt_emu {
krt_emu_f32x4 mfma_16x16x32_bf16(krt_emu_bf16x8 a, krt_emu_bf16x8 b, krt_emu_f32x4 c);
}
inline krt_emu_f32x4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(krt_emu_bf16x8 a, krt_emu_bf16x8 b, krt_emu_f32x4 c, int,
                                                             int, int) {
    return krt_emu::mfma_16x16x32_bf16(a, b, c);
}

// ------------------------------------------------------------- atomics
template <class T>
inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline float atomicAdd(float *p, float v) {
    float old = *p, want;
    do {
        want = old + v;
    } while (!__atomic_compare_exchange(p, &old, &want, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    return old;
}
template <class T>
inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicMin(T *p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
template <class T>
inline T atomicMax(T *p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}

// ------------------------------------------------------------- runtime API
typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorLaunchFailure = 719,
};
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum { hipStreamDefault = 0, hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
};
enum { hipHostMallocDefault = 0 };

struct krt_emu_stream;
struct krt_emu_event;
struct krt_emu_graph;
typedef krt_emu_stream *hipStream_t;
typedef krt_emu_event *hipEvent_t;
typedef krt_emu_graph *hipGraph_t;
typedef krt_emu_graph *hipGraphExec_t;
typedef void *hipGraphNode_t;

hipError_t hipMalloc(void **p, size_t n);
template <class T>
inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc(reinterpret_cast<void **>(p), n); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags = 0);
template <class T>
inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags = 0) {
    return hipHostMalloc(reinterpret_cast<void **>(p), n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void *dst, int v, size_t n);
hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s = nullptr);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t g, hipGraphNode_t *err_node, char *log,
                               size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes);
const char *hipGetErrorString(hipError_t e);

namespace krt_emu {
// The kernel call with its arguments evaluated and copied now.
template <class F, class... A>
std::function<void()> bind_args(F f, A... a) {
    return [f, tup = std::make_tuple(a...)]() mutable { std::apply(f, tup); };
}
// Runs grid x block lanes of `body` (the kernel call with its arguments bound
// by value), or records it when `s` is being captured into a graph.
void launch(const char *name, dim3 grid, dim3 block, size_t shmem, hipStream_t s, std::function<void()> body);
}  // namespace krt_emu

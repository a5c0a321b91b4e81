// Shadow-ray lists: kernels that resample (ReSTIR's spatial passes) append
// the visibility rays they need to a list, one register-lean kernel walks the
// list (any hit), and the next kernel reads one occlusion byte per ray slot.
// The walks then run at eight waves per SIMD instead of inside the shading
// code's register frame.
#pragma once

#include <hip/hip_runtime.h>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {
namespace {

struct RayList {
    float4 *rays;     // {origin, tmax}, {direction, slot as uint bits}
    uint32_t *count;  // rays appended (zeroed before the producing kernel)
};

// Wave-aggregated append: one atomic per wavefront; every lane of the wave
// must call it (lanes without a ray pass valid = false).
KRT_DEV void emit_ray(const RayList &L, bool valid, f3 o, f3 d, float tmax, uint32_t slot) {
    const unsigned long long want = __ballot(valid);
    if (!want) return;
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __ffsll((unsigned long long)want) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(L.count, (uint32_t)__popcll(want));
    base = __shfl(base, leader, 64);
    if (valid) {
        const uint32_t k = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
        L.rays[2 * (size_t)k] = make_float4(o.x, o.y, o.z, tmax);
        L.rays[2 * (size_t)k + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(slot));
    }
}

// Any-hit walks of a ray list (occluded(), amadeus_common.cuh:467-478): the
// list's count is read on the device; grid-stride over the rays.
template <bool INST>
__global__ __launch_bounds__(kTraceBlock) void shadow_list_kernel(SceneView sv, const float4 *__restrict__ rays,
                                                                  const uint32_t *__restrict__ count,
                                                                  uint8_t *__restrict__ occ, int *spill,
                                                                  unsigned long long *ctr) {
    __shared__ int lds_stack[kLdsDepthLean * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepthLean, spill + gtid, (int)nthreads};
    TravStats ts;
    unsigned nr = 0;
    const uint32_t n = *count;
    for (uint32_t k = gtid; k < n; k += nthreads) {
        const float4 a = rays[2 * (size_t)k], b = rays[2 * (size_t)k + 1];
        const bool r = occluded<true, INST>(sv, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, stk, ts);
        occ[__float_as_uint(b.w)] = r ? 1 : 0;
        nr++;
    }
    wave_count(ctr ? ctr + 3 : nullptr, nr);
    wave_count(ctr ? ctr + 6 : nullptr, ts.nodes);
    wave_count(ctr ? ctr + 9 : nullptr, ts.tris);
}

// Walks a list on the context stream with `blocks` workgroups of kTraceBlock
// lanes, whose spill columns the caller sized for kLdsDepthLean.
inline void walk_shadow_list(Context &ctx, bool instanced, const SceneView &sv, const RayList &L, uint8_t *occ,
                             int *spill, unsigned long long *ctr, uint32_t blocks) {
    if (instanced)
        shadow_list_kernel<true><<<blocks, kTraceBlock, 0, ctx.stream>>>(sv, L.rays, L.count, occ, spill, ctr);
    else
        shadow_list_kernel<false><<<blocks, kTraceBlock, 0, ctx.stream>>>(sv, L.rays, L.count, occ, spill, ctr);
    KRT_HIP(hipGetLastError());
}

}  // namespace
}  // namespace krt

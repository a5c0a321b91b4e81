// Shadow-ray lists: kernels that resample (ReSTIR's spatial passes) append
// the visibility rays they need to a list, one register-lean kernel walks the
// list (any hit), and the next kernel reads one occlusion byte per ray slot.
// The walks then run at eight waves per SIMD instead of inside the shading
// code's register frame.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {
namespace {

// Producers append with no global atomics: workgroup b of a producing launch
// (numbered by xcd_block(), the order in which its tiles lie in the image)
// owns the segment [b * seg, (b + 1) * seg) of `rays` and counts its rays in
// LDS; counts[b] is written by every producing workgroup, so nothing needs
// clearing between launches.  Before the walk, the segments are packed into
// one dense list (`dense`, offsets from an exclusive scan of the counts into
// `prefix`), so every walking lane gets the same share of rays.
struct RayList {
    float4 *rays;      // per ray {origin, tmax}, {direction, slot as uint bits}
    uint32_t *counts;  // rays in each segment
    uint32_t seg;      // slots per segment (lanes per workgroup x rays per lane)
    uint32_t nseg;     // producing workgroups
    uint32_t *prefix;  // nseg + 1: segment offsets in the dense list, total last
    float4 *dense;
};

// Ray records are read once per walk: streaming loads keep them from
// displacing tree nodes in L2.
KRT_DEV float4 stream_load(const float4 *p) {
    const float *f = reinterpret_cast<const float *>(p);
    return make_float4(__builtin_nontemporal_load(f), __builtin_nontemporal_load(f + 1),
                       __builtin_nontemporal_load(f + 2), __builtin_nontemporal_load(f + 3));
}

// Wave-aggregated append to the workgroup's segment: one LDS atomic per
// wavefront.  Every lane of the wave must call it (lanes without a ray pass
// valid = false); `nray` is the workgroup's LDS counter, zeroed before the
// first call and published with publish_rays after the last.
KRT_DEV void emit_ray(const RayList &L, uint32_t *nray, bool valid, f3 o, f3 d, float tmax, uint32_t slot) {
    const unsigned long long want = __ballot(valid);
    if (!want) return;
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __ffsll((unsigned long long)want) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(nray, (uint32_t)__popcll(want));
    base = __shfl(base, leader, 64);
    if (valid) {
        const size_t k = (size_t)xcd_block() * L.seg + base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
        L.rays[2 * k] = make_float4(o.x, o.y, o.z, tmax);
        L.rays[2 * k + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(slot));
    }
}

KRT_DEV void publish_rays(const RayList &L, const uint32_t *nray) {
    __syncthreads();
    if (threadIdx.x == 0) L.counts[xcd_block()] = *nray;
}

constexpr int kScanBlock = 1024;

// Exclusive scan of the segment counts in one workgroup: each lane sums a run
// of consecutive segments, the run sums are scanned in LDS, the lanes write
// their runs' offsets; prefix[nseg] = total.
__global__ __launch_bounds__(kScanBlock) void seg_scan_kernel(RayList L) {
    __shared__ uint32_t part[kScanBlock];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (L.nseg + kScanBlock - 1) / kScanBlock;
    const uint32_t b0 = min(t * per, L.nseg), b1 = min(b0 + per, L.nseg);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += L.counts[b];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kScanBlock; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (uint32_t b = b0; b < b1; b++) {
        L.prefix[b] = run;
        run += L.counts[b];
    }
    if (t == kScanBlock - 1) L.prefix[L.nseg] = part[kScanBlock - 1];
}

constexpr int kCompactBlock = 256;

// Direction class of a shadow ray: its octahedral projection on a
// kDirCells x kDirCells grid, so that rays of a segment toward nearby light
// samples share a class.
constexpr int kDirCells = 8;
KRT_DEV uint32_t dir_class(float x, float y, float z) {
    const float s = fabsf(x) + fabsf(y) + fabsf(z);
    float u = s > 0.0f ? x / s : 0.0f, v = s > 0.0f ? y / s : 0.0f;
    if (z < 0.0f) {
        const float uu = (1.0f - fabsf(v)) * (u < 0.0f ? -1.0f : 1.0f);
        v = (1.0f - fabsf(u)) * (v < 0.0f ? -1.0f : 1.0f);
        u = uu;
    }
    const int iu = min(kDirCells - 1, max(0, (int)((u * 0.5f + 0.5f) * (float)kDirCells)));
    const int iv = min(kDirCells - 1, max(0, (int)((v * 0.5f + 0.5f) * (float)kDirCells)));
    return (uint32_t)(iv * kDirCells + iu);
}

// Segment b -> dense[prefix[b] ...], one workgroup per segment.  Shadow rays
// (`by_octant`) are grouped by direction class with a counting sort in LDS,
// so that a wavefront of the walk starts rays of similar directions (their
// targets are scattered light samples); closest rays keep their order, in
// which a wavefront's rays come from one 8x8 tile.  Results are written by
// slot, so the order changes no answer.
__global__ __launch_bounds__(kCompactBlock) void seg_compact_kernel(RayList L, bool by_octant) {
    constexpr int NB = kDirCells * kDirCells;
    __shared__ uint32_t bins[NB], base[NB];
    const uint32_t b = blockIdx.x;
    const uint32_t n = L.counts[b];
    const float4 *src = L.rays + 2 * (size_t)b * L.seg;
    float4 *dst = L.dense + 2 * (size_t)L.prefix[b];
    if (!by_octant) {
        for (uint32_t k = threadIdx.x; k < 2 * n; k += kCompactBlock) dst[k] = src[k];
        return;
    }
    if (threadIdx.x < NB) bins[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += kCompactBlock) {
        const float4 c = src[2 * k + 1];
        atomicAdd(&bins[dir_class(c.x, c.y, c.z)], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int o = 0; o < NB; o++) {
            base[o] = run;
            run += bins[o];
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += kCompactBlock) {
        const float4 a = src[2 * k], c = src[2 * k + 1];
        const uint32_t at = atomicAdd(&base[dir_class(c.x, c.y, c.z)], 1u);
        dst[2 * at] = a;
        dst[2 * at + 1] = c;
    }
}

// Any-hit walks of the dense list (occluded(), amadeus_common.cuh:467-478),
// grid-stride; the count is read on the device.  Rays, node visits and
// triangle tests are added to the pass's counters (ctr[1 + pass],
// ctr[4 + pass], ctr[7 + pass]).
template <bool INST, bool WIDE>
__global__ __launch_bounds__(kTraceBlock) void shadow_list_kernel(SceneView sv, RayList L, uint8_t *__restrict__ occ,
                                                                  int *spill, unsigned long long *ctr, int pass) {
    __shared__ int lds_stack[kLdsDepthLean * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepthLean, spill + gtid, (int)nthreads};
    TravStats ts;
    unsigned nr = 0;
    const uint32_t n = L.prefix[L.nseg];
    for (uint32_t k = gtid; k < n; k += nthreads) {
        const float4 a = stream_load(L.dense + 2 * (size_t)k), c = stream_load(L.dense + 2 * (size_t)k + 1);
        const bool r = occluded<WIDE, INST>(sv, mk(a.x, a.y, a.z), mk(c.x, c.y, c.z), a.w, stk, ts);
        __builtin_nontemporal_store((uint8_t)(r ? 1 : 0), occ + __float_as_uint(c.w));
        nr++;
    }
    wave_count(ctr ? ctr + 1 + pass : nullptr, nr);
    wave_count(ctr ? ctr + 4 + pass : nullptr, ts.nodes);
    wave_count(ctr ? ctr + 7 + pass : nullptr, ts.tris);
}

// Closest-hit walks of the dense list (the binary tree, as the fused kernels
// walk it): hits[slot] = {t, b1, b2, prim bits}, hit_inst[slot] = instance
// (two-level scenes).  LDS stack depth D: 16 entries keep eight waves per
// SIMD resident, 32 keep five with fewer spills.
template <bool INST, int D>
__global__ __launch_bounds__(kTraceBlock) void closest_list_kernel(SceneView sv, RayList L, float4 *__restrict__ hits,
                                                                   int *__restrict__ hit_inst, int *spill,
                                                                   unsigned long long *ctr, int pass) {
    __shared__ int lds_stack[D * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, D, spill + gtid, (int)nthreads};
    TravStats ts;
    unsigned nr = 0;
    const uint32_t n = L.prefix[L.nseg];
    for (uint32_t k = gtid; k < n; k += nthreads) {
        const float4 a = stream_load(L.dense + 2 * (size_t)k), c = stream_load(L.dense + 2 * (size_t)k + 1);
        Hit h;
        trace_closest<INST>(sv, mk(a.x, a.y, a.z), mk(c.x, c.y, c.z), a.w, stk, h, ts);
        const uint32_t slot = __float_as_uint(c.w);
        float *hp = reinterpret_cast<float *>(hits + slot);
        __builtin_nontemporal_store(h.t, hp);
        __builtin_nontemporal_store(h.b1, hp + 1);
        __builtin_nontemporal_store(h.b2, hp + 2);
        __builtin_nontemporal_store(__int_as_float(h.prim), hp + 3);
        if (INST) hit_inst[slot] = h.inst;
        nr++;
    }
    wave_count(ctr, nr);
    wave_count(ctr ? ctr + 4 + pass : nullptr, ts.nodes);
    wave_count(ctr ? ctr + 7 + pass : nullptr, ts.tris);
}

// Workgroups of the walk: persistent grid over the list's capacity.
inline uint32_t walk_blocks(const RayList &L) {
    const size_t cap = (size_t)L.nseg * L.seg;
    return persistent_blocks((uint32_t)std::min<size_t>(cap, 0xffffffffu), kTraceBlock, kMaxResidentBlocks);
}

// Packs and walks a list on the context stream (the 4-wide tree unless
// `wide` is false); the caller sized the spill slab for kMaxResidentBlocks *
// kTraceBlock lanes at kLdsDepthLean.
inline void walk_shadow_list(Context &ctx, bool instanced, const SceneView &sv, const RayList &L, uint8_t *occ,
                             int *spill, unsigned long long *ctr, int pass, bool wide = true) {
    if (L.nseg == 0) return;
    seg_scan_kernel<<<1, kScanBlock, 0, ctx.stream>>>(L);
    KRT_HIP(hipGetLastError());
    seg_compact_kernel<<<L.nseg, kCompactBlock, 0, ctx.stream>>>(L, true);
    KRT_HIP(hipGetLastError());
    const uint32_t nb = walk_blocks(L);
    if (instanced && wide)
        shadow_list_kernel<true, true><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, occ, spill, ctr, pass);
    else if (instanced)
        shadow_list_kernel<true, false><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, occ, spill, ctr, pass);
    else if (wide)
        shadow_list_kernel<false, true><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, occ, spill, ctr, pass);
    else
        shadow_list_kernel<false, false><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, occ, spill, ctr, pass);
    KRT_HIP(hipGetLastError());
}

// Packs a list, then walks it for closest hits on the context stream (spill
// slab sized for kMaxResidentBlocks * kTraceBlock lanes at kLdsDepthLean).
inline void walk_closest_list(Context &ctx, bool instanced, const SceneView &sv, const RayList &L, float4 *hits,
                              int *hit_inst, int *spill, unsigned long long *ctr, int pass) {
    if (L.nseg == 0) return;
    seg_scan_kernel<<<1, kScanBlock, 0, ctx.stream>>>(L);
    KRT_HIP(hipGetLastError());
    seg_compact_kernel<<<L.nseg, kCompactBlock, 0, ctx.stream>>>(L, false);
    KRT_HIP(hipGetLastError());
    const uint32_t nb = walk_blocks(L);
    if (instanced)
        closest_list_kernel<true, kLdsDepthLean><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, hits, hit_inst, spill, ctr,
                                                                                     pass);
    else
        closest_list_kernel<false, kLdsDepthLean><<<nb, kTraceBlock, 0, ctx.stream>>>(sv, L, hits, hit_inst, spill,
                                                                                      ctr, pass);
    KRT_HIP(hipGetLastError());
}

// One launch for a level's two walks (the shadow rays of level L and the
// closest rays of level L + 1, both produced by the level-L shading kernel):
// the first `nbc` workgroups walk the closest list, the rest the shadow list
// (4-wide tree), so the two walks' tails overlap and one launch gap goes.
// Spill columns are indexed over the whole grid.
template <bool INST>
__global__ __launch_bounds__(kTraceBlock) void walk_pair_kernel(SceneView sv, RayList cl, float4 *__restrict__ hits,
                                                                int *__restrict__ hit_inst, RayList sl,
                                                                uint8_t *__restrict__ occ, int *spill,
                                                                unsigned long long *ctr, int pass, uint32_t nbc) {
    __shared__ int lds_stack[kLdsDepthLean * kTraceBlock];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepthLean, spill + gid, (int)(gridDim.x * blockDim.x)};
    TravStats ts;
    unsigned nr = 0;
    if (blockIdx.x < nbc) {
        const uint32_t nthreads = nbc * blockDim.x;
        const uint32_t n = cl.prefix[cl.nseg];
        for (uint32_t k = gid; k < n; k += nthreads) {
            const float4 a = stream_load(cl.dense + 2 * (size_t)k), c = stream_load(cl.dense + 2 * (size_t)k + 1);
            Hit h;
            trace_closest<INST>(sv, mk(a.x, a.y, a.z), mk(c.x, c.y, c.z), a.w, stk, h, ts);
            const uint32_t slot = __float_as_uint(c.w);
            float *hp = reinterpret_cast<float *>(hits + slot);
            __builtin_nontemporal_store(h.t, hp);
            __builtin_nontemporal_store(h.b1, hp + 1);
            __builtin_nontemporal_store(h.b2, hp + 2);
            __builtin_nontemporal_store(__int_as_float(h.prim), hp + 3);
            if (INST) hit_inst[slot] = h.inst;
            nr++;
        }
        wave_count(ctr, nr);
        wave_count(ctr ? ctr + 4 + pass : nullptr, ts.nodes);
        wave_count(ctr ? ctr + 7 + pass : nullptr, ts.tris);
    } else {
        const uint32_t nthreads = (gridDim.x - nbc) * blockDim.x;
        const uint32_t n = sl.prefix[sl.nseg];
        for (uint32_t k = gid - nbc * blockDim.x; k < n; k += nthreads) {
            const float4 a = stream_load(sl.dense + 2 * (size_t)k), c = stream_load(sl.dense + 2 * (size_t)k + 1);
            const bool r = occluded<true, INST>(sv, mk(a.x, a.y, a.z), mk(c.x, c.y, c.z), a.w, stk, ts);
            __builtin_nontemporal_store((uint8_t)(r ? 1 : 0), occ + __float_as_uint(c.w));
            nr++;
        }
        wave_count(ctr ? ctr + 1 + pass : nullptr, nr);
        wave_count(ctr ? ctr + 4 + pass : nullptr, ts.nodes);
        wave_count(ctr ? ctr + 7 + pass : nullptr, ts.tris);
    }
}

// Packs both lists, then walks them in one launch (walk_pair_kernel); the
// caller sized the spill slab for 2 * kMaxResidentBlocks * kTraceBlock lanes
// at kLdsDepthLean.
inline void walk_pair(Context &ctx, bool instanced, const SceneView &sv, const RayList &cl, float4 *hits,
                      int *hit_inst, const RayList &sl, uint8_t *occ, int *spill, unsigned long long *ctr, int pass) {
    if (cl.nseg == 0 || sl.nseg == 0) return;
    seg_scan_kernel<<<1, kScanBlock, 0, ctx.stream>>>(cl);
    KRT_HIP(hipGetLastError());
    seg_compact_kernel<<<cl.nseg, kCompactBlock, 0, ctx.stream>>>(cl, false);
    KRT_HIP(hipGetLastError());
    seg_scan_kernel<<<1, kScanBlock, 0, ctx.stream>>>(sl);
    KRT_HIP(hipGetLastError());
    seg_compact_kernel<<<sl.nseg, kCompactBlock, 0, ctx.stream>>>(sl, true);
    KRT_HIP(hipGetLastError());
    const uint32_t nbc = walk_blocks(cl), nbs = walk_blocks(sl);
    if (instanced)
        walk_pair_kernel<true><<<nbc + nbs, kTraceBlock, 0, ctx.stream>>>(sv, cl, hits, hit_inst, sl, occ, spill, ctr,
                                                                          pass, nbc);
    else
        walk_pair_kernel<false><<<nbc + nbs, kTraceBlock, 0, ctx.stream>>>(sv, cl, hits, hit_inst, sl, occ, spill,
                                                                           ctr, pass, nbc);
    KRT_HIP(hipGetLastError());
}

}  // namespace
}  // namespace krt

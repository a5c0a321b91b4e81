// Ray-query kernels behind krt_trace_closest / krt_trace_occluded.
#include <hip/hip_runtime.h>

#include <mutex>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {

namespace {

template <bool ANY, bool INST>
__global__ __launch_bounds__(kTraceBlock) void trace_kernel(const float4 *__restrict__ nodes,
                                                             const float4 *__restrict__ wide,
                                                             const float4 *__restrict__ tri,
                                                             const float4 *__restrict__ tlas,
                                                             const float4 *__restrict__ inst,
                                                             const float4 *__restrict__ rays, uint32_t n,
                                                             float4 *__restrict__ hits, int32_t *__restrict__ occ,
                                                             int *spill, int spill_depth) {
    __shared__ int lds_stack[kLdsDepth * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepth, spill + gtid, (int)nthreads};
    (void)spill_depth;
    TravStats ts;
    for (uint32_t i = gtid; i < n; i += nthreads) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        Hit h;
        bool r = INST ? traverse_inst<ANY>(tlas, inst, wide, tri, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, stk, h, ts)
                      : traverse<ANY>(nodes, wide, tri, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, stk, h, ts);
        if (ANY) {
            occ[i] = r ? 1 : 0;
        } else {
            hits[i] = make_float4(h.t, h.b1, h.b2, __int_as_float(h.prim));
        }
    }
}

__global__ void camera_rays_kernel(LaunchInfo li, int jitter, float4 *rays) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= li.width * li.height) return;
    uint32_t px = pix % li.width, py = pix / li.width;
    f3 seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
    f3 o, d;
    make_ray(li, px, py, seed, jitter != 0, o, d);
    rays[2 * pix] = make_float4(o.x, o.y, o.z, kTraceTmax);
    rays[2 * pix + 1] = make_float4(d.x, d.y, d.z, 0.0f);
}

// Plain streaming copy (HBM bandwidth and PMC byte-counter calibration).
__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

float copy_probe(Context &ctx, size_t bytes, int reps) {
    size_t n = bytes / sizeof(float4);
    DevBuf a(n * sizeof(float4)), b(n * sizeof(float4));
    KRT_HIP(hipMemsetAsync(a.p, 0, a.bytes, ctx.stream));
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    copy_kernel<<<256 * 8, 256, 0, ctx.stream>>>(a.as<float4>(), b.as<float4>(), n);  // warm
    KRT_HIP(hipEventRecord(e0, ctx.stream));
    for (int r = 0; r < reps; r++) copy_kernel<<<256 * 8, 256, 0, ctx.stream>>>(a.as<float4>(), b.as<float4>(), n);
    KRT_HIP(hipEventRecord(e1, ctx.stream));
    KRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / (float)reps;
}

void camera_rays_dev(Context &ctx, const LaunchInfo &li, int jitter, float *rays) {
    uint32_t n = li.width * li.height;
    camera_rays_kernel<<<(n + 255) / 256, 256, 0, ctx.stream>>>(li, jitter, (float4 *)rays);
    KRT_HIP(hipGetLastError());
}

int *spill_buffer(Context &ctx, Scene &s, uint32_t threads, int lds_depth) {
    size_t depth = s.accel.depth_bound > (uint32_t)lds_depth ? s.accel.depth_bound - lds_depth + 1 : 1;
    size_t need = sizeof(int) * depth * threads;
    if (ctx.spill.bytes < need) ctx.spill.alloc(need);
    return ctx.spill.as<int>();
}

uint32_t persistent_blocks(uint32_t work, uint32_t block, uint32_t max_blocks) {
    uint32_t b = (work + block - 1) / block;
    if (b < 1) b = 1;
    return b < max_blocks ? b : max_blocks;
}

static void trace_common(Context &ctx, Scene &s, const float *rays_dev, uint32_t n, float *hits_dev,
                         int32_t *occ_dev, bool any) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure (call krt_accel_build)");
    KRT_REQUIRE(!s.stale, KRT_ERR_STATE, "vertices moved: call krt_accel_refit or krt_accel_build first");
    if (n == 0) return;
    uint32_t nb = persistent_blocks(n, kTraceBlock, kMaxResidentBlocks);
    int *spill = spill_buffer(ctx, s, nb * kTraceBlock);
    const float4 *nodes = s.accel.nodes.as<float4>(), *wide = s.accel.wide.as<float4>(),
                 *tri = s.accel.tri.as<float4>(), *tlas = s.tlas.wide.as<float4>(), *inst = s.inst_rec.as<float4>();
    const float4 *rays = (const float4 *)rays_dev;
    const int depth = (int)s.accel.depth_bound;
    if (any && s.instanced)
        trace_kernel<true, true><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst, rays, n, nullptr,
                                                                    occ_dev, spill, depth);
    else if (any)
        trace_kernel<true, false><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst, rays, n, nullptr,
                                                                     occ_dev, spill, depth);
    else if (s.instanced)
        trace_kernel<false, true><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst, rays, n,
                                                                     (float4 *)hits_dev, nullptr, spill, depth);
    else
        trace_kernel<false, false><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst, rays, n,
                                                                      (float4 *)hits_dev, nullptr, spill, depth);
    KRT_HIP(hipGetLastError());
}

void trace_closest_dev(Context &ctx, Scene &s, const float *rays, uint32_t n, float *hits) {
    trace_common(ctx, s, rays, n, hits, nullptr, false);
}

// Traversal-stack overflow slab of trace_closest_dev_noalloc.  In host code
// `scratch` names a host-side stand-in, not device memory: kernels get the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int scratch[kNoAllocSpillInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int *noalloc_scratch() {
    int dev = 0;
    KRT_HIP(hipGetDevice(&dev));
    KRT_REQUIRE(dev >= 0 && dev < kMaxDevices, KRT_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        KRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        KRT_REQUIRE(p != nullptr, KRT_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch slab");
        g_scratch_ptr[dev] = (int *)p;
    }
    return g_scratch_ptr[dev];
}

// Same kernel and results as trace_closest_dev; the persistent grid is cut to
// what the slab holds (depth x 128 lanes per workgroup).  Nothing is
// allocated and nothing needs clearing: a stack entry is written before it is
// read.  Calls that overlap on one device share the slab and race.
void trace_closest_dev_noalloc(Context &ctx, Scene &s, const float *rays_dev, uint32_t n, float *hits_dev) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure (call krt_accel_build)");
    KRT_REQUIRE(!s.stale, KRT_ERR_STATE, "vertices moved: call krt_accel_refit or krt_accel_build first");
    KRT_REQUIRE(ctx.noalloc_spill != nullptr, KRT_ERR_STATE, "context has no scratch slab address");
    const size_t depth = s.accel.depth_bound > (uint32_t)kLdsDepth ? s.accel.depth_bound - kLdsDepth + 1 : 1;
    const size_t per_block = depth * (size_t)kTraceBlock;
    KRT_REQUIRE(per_block <= kNoAllocSpillInts, KRT_ERR_BUDGET,
                "traversal stack of this tree (depth bound " + std::to_string(s.accel.depth_bound) +
                    ") exceeds the no-allocation slab of " + std::to_string(kNoAllocSpillInts) + " ints");
    if (n == 0) return;
    const uint32_t cap_blocks = (uint32_t)(kNoAllocSpillInts / per_block);
    uint32_t nb = persistent_blocks(n, kTraceBlock, kMaxResidentBlocks);
    if (nb > cap_blocks) nb = cap_blocks;
    const float4 *nodes = s.accel.nodes.as<float4>(), *wide = s.accel.wide.as<float4>(),
                 *tri = s.accel.tri.as<float4>(), *tlas = s.tlas.wide.as<float4>(), *inst = s.inst_rec.as<float4>();
    const int dbound = (int)s.accel.depth_bound;
    if (s.instanced)
        trace_kernel<false, true><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst,
                                                                     (const float4 *)rays_dev, n, (float4 *)hits_dev,
                                                                     nullptr, ctx.noalloc_spill, dbound);
    else
        trace_kernel<false, false><<<nb, kTraceBlock, 0, ctx.stream>>>(nodes, wide, tri, tlas, inst,
                                                                      (const float4 *)rays_dev, n, (float4 *)hits_dev,
                                                                      nullptr, ctx.noalloc_spill, dbound);
    KRT_HIP(hipGetLastError());
}

void trace_occluded_dev(Context &ctx, Scene &s, const float *rays, uint32_t n, int32_t *occ) {
    trace_common(ctx, s, rays, n, nullptr, occ, true);
}

}  // namespace krt

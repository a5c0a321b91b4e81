// Path tracer attachment kernel: one pixel sample per lane over a persistent
// grid, the estimator of pt_pixel (krt_pt.h) written out in the kernel: the
// inlined call compiles to a 30 % slower loop here (same registers, different
// block layout), and the two are held equal bit for bit by the serving tests.
#include <hip/hip_runtime.h>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {

namespace {

struct PtArgs {
    LaunchInfo li;
    SceneView sv;
    float4 *hist;
    int *spill;
    unsigned long long *ctr;  // [0] closest, [1] shadow
    double *acc64;            // optional FP64 sums of samples, then of their squares (3 + 3 per pixel)
    uint32_t *work;           // work-item counter of the launch (zeroed before it)
};

// Path regeneration on a persistent grid: a lane whose path has ended takes
// the next pixel at once instead of idling until the longest path of its
// wavefront ends.  Lanes that need work are counted with a wave ballot and
// the wave takes that many consecutive work items with one atomic, so the
// pixels a wave starts together come from one 8x8 tile.  Every pixel's
// estimate is computed exactly as before; only its lane and start time
// change.
template <bool INST>
__global__ __launch_bounds__(kTraceBlock) void pt_kernel(PtArgs A) {
    __shared__ int lds_stack[kLdsDepth * kTraceBlock];
    const LaunchInfo &li = A.li;
    const SceneView &sv = A.sv;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepth, A.spill + gtid, (int)nthreads};
    History hist{A.hist + gtid, nthreads};
    TravStats ts;
    RayCount rc;
    const float count = (float)li.samples;
    const TileOrder tiles(li.width, li.height);
    bool active = false, done = false;
    uint32_t pix = 0;
    int level = 0, nlev = 0;
    f3 seed = mk(0, 0, 0), o = mk(0, 0, 0), d = mk(0, 0, 0);
    f3 pos0 = mk(0, 0, 0), n0 = mk(0, 0, 0), alb0 = mk(0, 0, 0);
    for (;;) {
        // refill: one atomic per wavefront for all lanes without a path
        const bool need = !active && !done;
        const unsigned long long want = __ballot(need);
        if (want) {
            const int leader = __ffsll((unsigned long long)want) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(A.work, (uint32_t)__popcll(want));
            base = __shfl(base, leader, 64);
            if (need) {
                const uint32_t item = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
                uint32_t px, py;
                if (item >= tiles.items) {
                    done = true;
                } else if (tiles.pixel(item, px, py)) {
                    pix = py * li.width + px;
                    seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
                    make_ray(li, px, py, seed, li.jitter != 0, o, d);
                    pos0 = n0 = alb0 = mk(0, 0, 0);
                    level = 0;
                    nlev = 0;
                    active = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
        if (!active) continue;
        // one level of the path (amadeus_path_tracer.cu:133-201)
        bool end = true;
        Hit h;
        rc.closest++;
        trace_closest<INST>(sv, o, d, kTraceTmax, stk, h, ts);
        Surface sf;
        if (h.prim >= 0 && load_surface<INST>(sv, h, d, sf)) {  // a miss (no environment map) or a NaN normal end it
            Material mat = load_material(sv.mats, sf.mat_id);
            f3 n = sf.n;
            f3 x = offset_point(sf.x, n, mat.type);
            SurfaceHit sh{mat, sf.entering, n, -d, x};
            f3 direct = mat.emission + Ld<true, INST>(sv, sh, seed, stk, ts, rc);
            int out;
            f3 wi;
            float pdf;
            f3 f = eval_bsdf(sh, wi, pdf, out, seed);
            f3 T = f * fabsf(dot(wi, n)) / pdf;
            bool has_pdf = pdf > 0.0f;
            hist.put(level, direct, has_pdf, T);
            if (level == 0) {
                pos0 = x;
                n0 = n;
                alb0 = mat.diffuse;
            }
            nlev = level + 1;
            if (has_pdf && level < li.max_depth) {
                o = x;
                d = wi;
                level++;
                end = false;
            }
        }
        if (!end) continue;
        // the path is complete: fold its levels from the deepest up
        f3 value = mk(0, 0, 0);
        for (int k = nlev - 1; k >= 0; k--) {
            f3 dk, T;
            bool has_pdf;
            hist.get(k, dk, has_pdf, T);
            f3 v = dk;
            if (has_pdf) v = v + T * value;
            value = v;
        }
        float4 col = make_float4(value.x, value.y, value.z, 1.0f);
        if (isnan3(value)) col = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
        else if (A.acc64) {  // converged-reference accumulator (FP64 only here)
            double *sum = A.acc64 + 3 * (size_t)pix;
            double *sq = A.acc64 + 3 * ((size_t)li.width * li.height + pix);  // per-pixel sample variance
            const double v[3] = {(double)value.x, (double)value.y, (double)value.z};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sum[c] += v[c];
                sq[c] += v[c] * v[c];
            }
        }
        const bool acc = li.accumulate != 0;
        li.color[pix] = accumulate4(li.color[pix], col, acc, count);
        li.normal[pix] = accumulate4(li.normal[pix], make_float4(n0.x, n0.y, n0.z, 0.0f), acc, count);
        li.albedo[pix] = accumulate4(li.albedo[pix], make_float4(alb0.x, alb0.y, alb0.z, 0.0f), acc, count);
        li.position[pix] = make_float4(pos0.x, pos0.y, pos0.z, 0.0f);
        active = false;
    }
    wave_count(A.ctr, rc.closest);
    wave_count(A.ctr ? A.ctr + 1 : nullptr, rc.shadow);
    wave_count(A.ctr ? A.ctr + 4 : nullptr, ts.nodes);
    wave_count(A.ctr ? A.ctr + 7 : nullptr, ts.tris);
}

}  // namespace

void pt_render(Context &ctx, Scene &s, const LaunchInfo &li, DevBuf &history, unsigned long long *ctr,
               PassTimes &times, double *acc64) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(li.max_depth >= 0 && li.max_depth <= 64, KRT_ERR_INVALID, "max_depth must be in [0, 64]");
    const uint32_t npix = li.width * li.height;
    uint32_t nb = persistent_blocks(npix, kTraceBlock, kMaxResidentBlocks);
    uint32_t nthreads = nb * kTraceBlock;
    size_t need = sizeof(float4) * 2 * (size_t)(li.max_depth + 1) * nthreads;
    if (history.bytes < need) history.alloc(need);
    uint32_t *work = ctx.counters.as<uint32_t>();
    KRT_HIP(hipMemsetAsync(work, 0, sizeof(uint32_t), ctx.stream));
    PtArgs A{li, make_view(s), history.as<float4>(), spill_buffer(ctx, s, nthreads), ctr, acc64, work};
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, ctx.stream));
    if (s.instanced) pt_kernel<true><<<nb, kTraceBlock, 0, ctx.stream>>>(A);
    else pt_kernel<false><<<nb, kTraceBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(e1, ctx.stream));
    KRT_HIP(hipEventSynchronize(e1));
    KRT_HIP(hipEventElapsedTime(&times.ms[0], e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace krt

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
};

template <bool INST>
__global__ __launch_bounds__(kTraceBlock) void pt_kernel(PtArgs A) {
    __shared__ int lds_stack[kLdsDepth * kTraceBlock];
    const LaunchInfo &li = A.li;
    const SceneView &sv = A.sv;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepth, A.spill + gtid, (int)nthreads};
    History hist{A.hist + gtid, nthreads};
    TravStats ts;
    RayCount rc;
    const float count = (float)li.samples;
    const TileOrder tiles(li.width, li.height);
    for (uint32_t item = gtid; item < tiles.items; item += nthreads) {
        uint32_t px, py;
        if (!tiles.pixel(item, px, py)) continue;
        const uint32_t pix = py * li.width + px;
        f3 seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
        f3 o, d;
        make_ray(li, px, py, seed, li.jitter != 0, o, d);
        f3 pos0 = mk(0, 0, 0), n0 = mk(0, 0, 0), alb0 = mk(0, 0, 0);
        int nlev = 0;
        for (int level = 0; level <= li.max_depth; level++) {
            Hit h;
            rc.closest++;
            trace_closest<INST>(sv, o, d, kTraceTmax, stk, h, ts);
            if (h.prim < 0) break;  // __miss__: no environment map, value 0
            Surface sf;
            if (!load_surface<INST>(sv, h, d, sf)) break;
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
            if (!has_pdf) break;
            o = x;
            d = wi;
        }
        f3 value = mk(0, 0, 0);
        for (int k = nlev - 1; k >= 0; k--) {
            f3 direct, T;
            bool has_pdf;
            hist.get(k, direct, has_pdf, T);
            f3 v = direct;
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
    PtArgs A{li, make_view(s), history.as<float4>(), spill_buffer(ctx, s, nthreads), ctr, acc64};
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

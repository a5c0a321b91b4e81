// Path tracer attachment kernel: one pixel sample per lane (pt_pixel,
// krt_pt.h) over a persistent grid.
#include <hip/hip_runtime.h>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_pt.h"
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
    double *acc64;            // optional FP64 running sum of samples (3 per pixel)
};

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
    const uint32_t npix = li.width * li.height;
    for (uint32_t pix = gtid; pix < npix; pix += nthreads)
        pt_pixel(li, sv, pix, li.samples, stk, hist, ts, rc, A.acc64);
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
    pt_kernel<<<nb, kTraceBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(e1, ctx.stream));
    KRT_HIP(hipEventSynchronize(e1));
    KRT_HIP(hipEventElapsedTime(&times.ms[0], e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace krt

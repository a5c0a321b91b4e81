// Path tracer with an online-trained radiance cache.
//
// Every pixel runs the path tracer's estimator (the passes of krt_pt.hip,
// after source/optix/amadeus_path_tracer.cu:73-228) up to the cache vertex,
// the hit at level `depth`; there the path stops and records a query {x, -d,
// n, albedo, roughness} for the MLP, whose prediction stands in for the
// radiance the rest of the path would have scattered toward the previous
// vertex:
//   value_0 = P + Theta * (Le(x_depth) + max(cache(x_depth, -d), 0)),
// with (P, Theta) folded from the levels before the cache vertex in the
// estimator's order.  The vertex's own emission is added exactly rather than
// learned: a light's edge is a step the encoding cannot resolve.  A strided
// subset of pixels (a new offset every frame) traces the full path instead,
// shows that unbiased value, and hands the value at its cache vertex less that
// vertex's emission to the trainer as the target for the same query.  Per
// frame: path-tracer passes -> MLP inference -> composite -> compaction of the
// training slots -> one optimiser step; nothing waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "krt_cache.h"
#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"

namespace krt {

namespace {

__global__ void composite_kernel(LaunchInfo li, const float4 *__restrict__ prefix, const float4 *__restrict__ theta,
                                 const float *__restrict__ pred) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= li.width * li.height) return;
    const float4 p = prefix[pix];
    f3 value = mk(p.x, p.y, p.z);
    const float4 t = theta[pix];
    if (t.w != 0.0f) {
        const f3 y = mk(fmaxf(pred[3 * (size_t)pix], 0.0f), fmaxf(pred[3 * (size_t)pix + 1], 0.0f),
                        fmaxf(pred[3 * (size_t)pix + 2], 0.0f));
        value = value + mk(t.x, t.y, t.z) * y;
    }
    float4 col = make_float4(value.x, value.y, value.z, 1.0f);
    if (isnan3(value)) col = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    li.color[pix] = accumulate4(li.color[pix], col, li.accumulate != 0, (float)li.samples);
}

__global__ void fold_kernel(uint32_t npix, const float4 *__restrict__ theta, const float *__restrict__ pred,
                            float4 *__restrict__ intermediate) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const float4 t = theta[pix];
    if (t.w == 0.0f) return;
    float4 im = intermediate[pix];
    im.x = im.x + t.x * fmaxf(pred[3 * (size_t)pix], 0.0f);
    im.y = im.y + t.y * fmaxf(pred[3 * (size_t)pix + 1], 0.0f);
    im.z = im.z + t.z * fmaxf(pred[3 * (size_t)pix + 2], 0.0f);
    if (isnan(im.x) || isnan(im.y) || isnan(im.z)) im = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    intermediate[pix] = im;
}

// Order-preserving compaction of the valid training slots, in chunks of
// kChunk slots: chunk_count_kernel counts each chunk's valid slots,
// chunk_scan_kernel turns the counts into the chunks' first batch rows (and
// the batch size), chunk_scatter_kernel ranks each chunk's valid slots and
// copies their records, the chunk's threads sharing the copy word by word.
constexpr int kChunk = 256;
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kChunk) void chunk_count_kernel(const uint32_t *__restrict__ slot_ok, uint32_t M,
                                                             uint32_t *__restrict__ chunk_n) {
    const uint32_t i = blockIdx.x * kChunk + threadIdx.x;
    const int ok = i < M && slot_ok[i] != 0u;
    const int n = __syncthreads_count(ok);
    if (threadIdx.x == 0) chunk_n[blockIdx.x] = (uint32_t)n;
}

__global__ __launch_bounds__(kScanThreads) void chunk_scan_kernel(uint32_t *__restrict__ chunk_n, uint32_t nchunk,
                                                                  uint32_t *__restrict__ n_out) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nchunk + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = min(t * per, nchunk), b1 = min(b0 + per, nchunk);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += chunk_n[b];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive: rows before this thread's chunks
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t n = chunk_n[b];
        chunk_n[b] = run;
        run += n;
    }
    if (t == kScanThreads - 1) *n_out = part[t];
}

__global__ __launch_bounds__(kChunk) void chunk_scatter_kernel(const float *__restrict__ slot_raw,
                                                               const float *__restrict__ slot_tgt,
                                                               const uint32_t *__restrict__ slot_ok, uint32_t M,
                                                               const uint32_t *__restrict__ chunk_first,
                                                               float *__restrict__ raw, float *__restrict__ tgt) {
    constexpr int kWaves = kChunk / 64;
    __shared__ uint32_t wave_n[kWaves];
    __shared__ int dst[kChunk];  // batch row of each slot of the chunk, -1 when invalid
    const uint32_t base = blockIdx.x * kChunk;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t i = base + t;
    const bool ok = i < M && slot_ok[i] != 0u;
    const unsigned long long bal = __ballot(ok);
    if (lane == 0) wave_n[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t row = chunk_first[blockIdx.x] + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    for (uint32_t k = 0; k < w; k++) row += wave_n[k];
    dst[t] = ok ? (int)row : -1;
    __syncthreads();
    const uint32_t nslot = min((uint32_t)kChunk, M - base);
    for (uint32_t j = t; j < nslot * kMlpRaw; j += kChunk) {
        const int r = dst[j / kMlpRaw];
        if (r >= 0) raw[(size_t)kMlpRaw * r + j % kMlpRaw] = slot_raw[(size_t)kMlpRaw * base + j];
    }
    for (uint32_t j = t; j < nslot * 3; j += kChunk) {
        const int r = dst[j / 3];
        if (r >= 0) tgt[3 * (size_t)r + j % 3] = slot_tgt[3 * (size_t)base + j];
    }
}

uint32_t frame_offset(uint64_t frame, uint32_t seed, uint32_t stride) {
    uint64_t z = frame * 0x9E3779B97F4A7C15ull + seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z % stride);
}

}  // namespace

CacheView cache_begin(Context &ctx, Scene &s, const CacheOptions &opt, CacheState &cs, uint32_t npix,
                      uint32_t user_seed) {
    KRT_REQUIRE(opt.depth >= 0, KRT_ERR_INVALID, "cache_depth must be >= 0");
    KRT_REQUIRE(opt.train_samples >= 0, KRT_ERR_INVALID, "train_samples must be >= 0");
    if (!cs.ready) {
        mlp_init(ctx, cs.net, cs.cfg, 0xC0FFEEull ^ user_seed);
        cs.ready = true;
    }
    // positions mapped to [0, 1] by the scene's bounding cube
    float ext = 0.0f;
    for (int a = 0; a < 3; a++) ext = fmaxf(ext, s.bmax[a] - s.bmin[a]);
    ext = ext > 0.0f ? ext : 1.0f;
    for (int a = 0; a < 3; a++) {
        cs.net.enc.pmin[a] = s.bmin[a];
        cs.net.enc.pscale[a] = 1.0f / ext;
    }
    if (cs.npix != npix) {
        cs.query.alloc(sizeof(float) * kMlpRaw * (size_t)npix);
        cs.pred.alloc(sizeof(float) * 3 * (size_t)npix);
        cs.prefix.alloc(sizeof(float4) * (size_t)npix);
        cs.theta.alloc(sizeof(float4) * (size_t)npix);
        KRT_HIP(hipMemsetAsync(cs.query.p, 0, cs.query.bytes, ctx.stream));
        cs.npix = npix;
    }
    const uint32_t M = opt.training ? (uint32_t)std::min<uint64_t>((uint64_t)opt.train_samples, npix) : 0u;
    if (cs.slots < M || !cs.slot_ok.p) {
        const uint32_t cap = M ? M : 1u;
        cs.slot_raw.alloc(sizeof(float) * kMlpRaw * (size_t)cap);
        cs.slot_tgt.alloc(sizeof(float) * 3 * (size_t)cap);
        cs.slot_ok.alloc(sizeof(uint32_t) * (size_t)cap);
        cs.batch_raw.alloc(sizeof(float) * kMlpRaw * (size_t)cap);
        cs.batch_tgt.alloc(sizeof(float) * 3 * (size_t)cap);
        cs.batch_n.alloc(sizeof(uint32_t));
        cs.chunk_n.alloc(sizeof(uint32_t) * ((cap + kChunk - 1) / kChunk));
        cs.slots = cap;
    }
    CacheView cv;
    cv.enabled = 1;
    cv.depth = opt.depth;
    cv.learn_emission = opt.learn_emission;
    cv.query = cs.query.as<float>();
    cv.theta = cs.theta.as<float4>();
    cv.slot_raw = cs.slot_raw.as<float>();
    cv.slot_tgt = cs.slot_tgt.as<float>();
    cv.slot_ok = cs.slot_ok.as<uint32_t>();
    cv.M = M;
    cv.stride = M ? npix / M : 1u;
    cv.offset = M ? frame_offset(cs.frame, user_seed, cv.stride) : npix;
    cs.frame++;
    return cv;
}

void cache_infer(CacheState &cs, uint32_t npix, hipStream_t stream) {
    mlp_infer_dev(cs.net, cs.query.as<float>(), npix, cs.pred.as<float>(), stream);
}

void cache_fold_intermediate(CacheState &cs, uint32_t npix, float4 *intermediate, hipStream_t stream) {
    fold_kernel<<<(npix + 255) / 256, 256, 0, stream>>>(npix, cs.theta.as<float4>(), cs.pred.as<float>(),
                                                       intermediate);
    KRT_HIP(hipGetLastError());
}

void cache_train(Context &ctx, CacheState &cs, const CacheView &cv) {
    if (cv.M == 0) return;
    const uint32_t nchunk = (cv.M + kChunk - 1) / kChunk;
    uint32_t *chunk = cs.chunk_n.as<uint32_t>();
    chunk_count_kernel<<<nchunk, kChunk, 0, ctx.stream>>>(cs.slot_ok.as<uint32_t>(), cv.M, chunk);
    KRT_HIP(hipGetLastError());
    chunk_scan_kernel<<<1, kScanThreads, 0, ctx.stream>>>(chunk, nchunk, cs.batch_n.as<uint32_t>());
    KRT_HIP(hipGetLastError());
    chunk_scatter_kernel<<<nchunk, kChunk, 0, ctx.stream>>>(cs.slot_raw.as<float>(), cs.slot_tgt.as<float>(),
                                                           cs.slot_ok.as<uint32_t>(), cv.M, chunk,
                                                           cs.batch_raw.as<float>(), cs.batch_tgt.as<float>());
    KRT_HIP(hipGetLastError());
    mlp_gradient_dev(ctx, cs.net, cs.batch_raw.as<float>(), cs.batch_tgt.as<float>(), cv.M, cs.batch_n.as<uint32_t>());
    mlp_apply_dev(ctx, cs.net);
}

void cache_pt_render(Context &ctx, Scene &s, const LaunchInfo &li, const CacheOptions &opt, CacheState &cs,
                     unsigned long long *ctr, PassTimes &times) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "the path tracer's radiance cache renders flat scenes (krt_scene_upload)");
    KRT_REQUIRE(li.max_depth >= 0 && li.max_depth <= 64, KRT_ERR_INVALID, "max_depth must be in [0, 64]");
    const uint32_t npix = li.width * li.height;
    const CacheView cv = cache_begin(ctx, s, opt, cs, npix, li.user_seed);
    // The inference reads only the query records, all written by the
    // cache-vertex shading pass: it runs on the side stream while the main
    // stream walks the training paths' remaining levels, and the composite
    // waits for it.
    hipStream_t side = ctx.side ? ctx.side : ctx.stream;
    hipEvent_t ev[5], queried, inf0, inf1;
    for (auto &e : ev) KRT_HIP(hipEventCreate(&e));
    KRT_HIP(hipEventCreate(&queried));
    KRT_HIP(hipEventCreate(&inf0));
    KRT_HIP(hipEventCreate(&inf1));
    KRT_HIP(hipEventRecord(ev[0], ctx.stream));
    pt_wave(ctx, s, li, ctr, nullptr, &cv, cs.prefix.as<float4>(), queried);
    KRT_HIP(hipEventRecord(ev[1], ctx.stream));
    KRT_HIP(hipStreamWaitEvent(side, queried, 0));
    KRT_HIP(hipEventRecord(inf0, side));
    cache_infer(cs, npix, side);
    KRT_HIP(hipEventRecord(inf1, side));
    KRT_HIP(hipStreamWaitEvent(ctx.stream, inf1, 0));
    KRT_HIP(hipEventRecord(ev[2], ctx.stream));
    composite_kernel<<<(npix + 255) / 256, 256, 0, ctx.stream>>>(li, cs.prefix.as<float4>(), cs.theta.as<float4>(),
                                                                cs.pred.as<float>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(ev[3], ctx.stream));
    cache_train(ctx, cs, cv);
    KRT_HIP(hipEventRecord(ev[4], ctx.stream));
    KRT_HIP(hipEventSynchronize(ev[4]));
    KRT_HIP(hipEventElapsedTime(&times.ms[0], ev[0], ev[1]));
    KRT_HIP(hipEventElapsedTime(&times.ms[1], inf0, inf1));
    KRT_HIP(hipEventElapsedTime(&times.ms[2], ev[2], ev[3]));
    KRT_HIP(hipEventElapsedTime(&times.ms[3], ev[3], ev[4]));
    KRT_HIP(hipEventElapsedTime(&times.frame, ev[0], ev[4]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(queried);
    (void)hipEventDestroy(inf0);
    (void)hipEventDestroy(inf1);
}

}  // namespace krt

// Path tracer with an online-trained radiance cache.
//
// Every pixel runs the path tracer's estimator (the loop of krt_pt.hip, after
// source/optix/amadeus_path_tracer.cu:73-228) up to the cache vertex, the hit
// at level `depth`; there the path stops and records a query {x, -d, n,
// albedo, roughness} for the MLP, whose prediction stands in for the radiance
// the rest of the path would have scattered toward the previous vertex:
//   value_0 = P + Theta * (Le(x_depth) + max(cache(x_depth, -d), 0)),
// with (P, Theta) folded from the levels before the cache vertex in the
// estimator's order.  The vertex's own emission is added exactly rather than
// learned: a light's edge is a step the encoding cannot resolve.  A strided subset of pixels (a new offset every frame)
// traces the full path instead, shows that unbiased value, and hands the value
// at its cache vertex less that vertex's emission to the trainer as the target
// for the same query.  Per
// frame: path kernel -> MLP inference -> composite -> compaction of the
// training slots -> one optimiser step; nothing waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "krt_cache.h"
#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {

namespace {

struct CacheArgs {
    LaunchInfo li;
    SceneView sv;
    float4 *hist;
    int *spill;
    unsigned long long *ctr;
    float4 *prefix;  // npix: P of value_0 = P + Theta * cache
    CacheView cv;
    uint32_t *work;  // work-item counter of the launch (zeroed before it)
};

__global__ __launch_bounds__(kTraceBlock) void cache_path_kernel(CacheArgs A) {
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
    const CacheView &cv = A.cv;
    const bool acc = li.accumulate != 0;
    const uint32_t lane = threadIdx.x & 63u;
    // Path regeneration (as the path tracer's kernel): most paths stop at the
    // cache vertex while the strided training paths run to full depth, so a
    // lane whose path has ended takes the next pixel at once instead of
    // waiting for a training path of its wavefront.
    bool active = false, done = false, train = false, queried = false;
    uint32_t pix = 0, slot = 0;
    int level = 0, nlev = 0;
    f3 seed = mk(0, 0, 0), o = mk(0, 0, 0), d = mk(0, 0, 0), le_c = mk(0, 0, 0);
    f3 pos0 = mk(0, 0, 0), n0 = mk(0, 0, 0), alb0 = mk(0, 0, 0);
    for (;;) {
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
                    slot = 0;
                    train = cv.training(pix, slot);
                    seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
                    make_ray(li, px, py, seed, li.jitter != 0, o, d);
                    pos0 = n0 = alb0 = mk(0, 0, 0);
                    level = 0;
                    nlev = 0;
                    queried = false;
                    le_c = mk(0, 0, 0);  // emission at the cache vertex, added exactly
                    active = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
        if (!active) continue;
        bool end = true;
        Hit h;
        rc.closest++;
        traverse<false>(sv.nodes, sv.wide, sv.tri, o, d, kTraceTmax, stk, h, ts);
        Surface sf;
        if (h.prim >= 0 && load_surface(sv, h, d, sf)) {
            Material mat = load_material(sv.mats, sf.mat_id);
            bool cache_stop = false;
            if (level == cv.depth) {
                put_query(train ? cv.slot_raw + (size_t)kMlpRaw * slot : cv.query + (size_t)kMlpRaw * pix, sf.x, -d,
                          sf.n, mat);
                queried = true;
                le_c = cv.learn_emission ? mk(0, 0, 0) : mat.emission;
                cache_stop = !train;
            }
            if (!cache_stop) {
                f3 n = sf.n;
                f3 x = offset_point(sf.x, n, mat.type);
                SurfaceHit sh{mat, sf.entering, n, -d, x};
                f3 direct = mat.emission + Ld<false>(sv, sh, seed, stk, ts, rc);  // binary: occupancy
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
        }
        if (!end) continue;
        float4 pre, th = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (queried && !train) {
            // value_0 = P + Theta * cache, levels 0 .. depth-1 folded outside in
            f3 P = mk(0, 0, 0), Th = mk(1.0f, 1.0f, 1.0f);
            for (int k = nlev - 1; k >= 0; k--) {
                f3 dk, T;
                bool has_pdf;
                hist.get(k, dk, has_pdf, T);
                P = dk + T * P;
                Th = T * Th;
            }
            P = P + Th * le_c;
            pre = make_float4(P.x, P.y, P.z, 0.0f);
            th = make_float4(Th.x, Th.y, Th.z, 1.0f);
        } else {
            f3 value = mk(0, 0, 0), target = mk(0, 0, 0);
            for (int k = nlev - 1; k >= 0; k--) {
                f3 dk, T;
                bool has_pdf;
                hist.get(k, dk, has_pdf, T);
                f3 v = dk;
                if (has_pdf) v = v + T * value;
                value = v;
                if (k == cv.depth) target = value - le_c;  // scattered radiance only
            }
            pre = make_float4(value.x, value.y, value.z, 0.0f);
            if (train) {
                const bool ok = queried && nlev > cv.depth && finite3(target);
                float *t = cv.slot_tgt + 3 * (size_t)slot;
                t[0] = target.x;
                t[1] = target.y;
                t[2] = target.z;
                cv.slot_ok[slot] = ok ? 1u : 0u;
            }
        }
        A.prefix[pix] = pre;
        cv.theta[pix] = th;
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

// Order-preserving compaction of the valid training slots (one workgroup).
constexpr int kCompactThreads = 1024;
__global__ __launch_bounds__(kCompactThreads) void compact_kernel(const float *__restrict__ slot_raw,
                                                                  const float *__restrict__ slot_tgt,
                                                                  const uint32_t *__restrict__ slot_ok, uint32_t M,
                                                                  float *__restrict__ raw, float *__restrict__ tgt,
                                                                  uint32_t *n_out) {
    __shared__ uint32_t s[kCompactThreads];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (M + kCompactThreads - 1) / kCompactThreads;
    const uint32_t b = tid * per, e = min(b + per, M);
    uint32_t cnt = 0;
    for (uint32_t i = b; i < e; i++) cnt += slot_ok[i];
    s[tid] = cnt;
    __syncthreads();
    for (uint32_t off = 1; off < kCompactThreads; off <<= 1) {
        const uint32_t v = tid >= off ? s[tid - off] : 0u;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    uint32_t dst = s[tid] - cnt;
    for (uint32_t i = b; i < e; i++) {
        if (!slot_ok[i]) continue;
        for (int a = 0; a < kMlpRaw; a++) raw[(size_t)kMlpRaw * dst + a] = slot_raw[(size_t)kMlpRaw * i + a];
        for (int a = 0; a < 3; a++) tgt[3 * (size_t)dst + a] = slot_tgt[3 * (size_t)i + a];
        dst++;
    }
    if (tid == kCompactThreads - 1) *n_out = s[tid];
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

void cache_infer(Context &ctx, CacheState &cs, uint32_t npix) {
    mlp_infer_dev(ctx, cs.net, cs.query.as<float>(), npix, cs.pred.as<float>());
}

void cache_fold_intermediate(Context &ctx, CacheState &cs, uint32_t npix, float4 *intermediate) {
    fold_kernel<<<(npix + 255) / 256, 256, 0, ctx.stream>>>(npix, cs.theta.as<float4>(), cs.pred.as<float>(),
                                                           intermediate);
    KRT_HIP(hipGetLastError());
}

void cache_train(Context &ctx, CacheState &cs, const CacheView &cv) {
    if (cv.M == 0) return;
    compact_kernel<<<1, kCompactThreads, 0, ctx.stream>>>(cs.slot_raw.as<float>(), cs.slot_tgt.as<float>(),
                                                         cs.slot_ok.as<uint32_t>(), cv.M, cs.batch_raw.as<float>(),
                                                         cs.batch_tgt.as<float>(), cs.batch_n.as<uint32_t>());
    KRT_HIP(hipGetLastError());
    mlp_gradient_dev(ctx, cs.net, cs.batch_raw.as<float>(), cs.batch_tgt.as<float>(), cv.M, cs.batch_n.as<uint32_t>());
    mlp_apply_dev(ctx, cs.net);
}

void cache_pt_render(Context &ctx, Scene &s, const LaunchInfo &li, const CacheOptions &opt, CacheState &cs,
                     DevBuf &history, unsigned long long *ctr, PassTimes &times) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "the path tracer's radiance cache renders flat scenes (krt_scene_upload)");
    KRT_REQUIRE(li.max_depth >= 0 && li.max_depth <= 64, KRT_ERR_INVALID, "max_depth must be in [0, 64]");
    const uint32_t npix = li.width * li.height;
    const CacheView cv = cache_begin(ctx, s, opt, cs, npix, li.user_seed);
    uint32_t nb = persistent_blocks(npix, kTraceBlock, kMaxResidentBlocks);
    uint32_t nthreads = nb * kTraceBlock;
    size_t need = sizeof(float4) * 2 * (size_t)(li.max_depth + 1) * nthreads;
    if (history.bytes < need) history.alloc(need);
    uint32_t *work = ctx.counters.as<uint32_t>();
    KRT_HIP(hipMemsetAsync(work, 0, sizeof(uint32_t), ctx.stream));
    CacheArgs A{li, make_view(s), history.as<float4>(), spill_buffer(ctx, s, nthreads), ctr, cs.prefix.as<float4>(),
                cv, work};
    hipEvent_t ev[5];
    for (auto &e : ev) KRT_HIP(hipEventCreate(&e));
    KRT_HIP(hipEventRecord(ev[0], ctx.stream));
    cache_path_kernel<<<nb, kTraceBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(ev[1], ctx.stream));
    cache_infer(ctx, cs, npix);
    KRT_HIP(hipEventRecord(ev[2], ctx.stream));
    composite_kernel<<<(npix + 255) / 256, 256, 0, ctx.stream>>>(li, cs.prefix.as<float4>(), cs.theta.as<float4>(),
                                                                cs.pred.as<float>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(ev[3], ctx.stream));
    cache_train(ctx, cs, cv);
    KRT_HIP(hipEventRecord(ev[4], ctx.stream));
    KRT_HIP(hipEventSynchronize(ev[4]));
    for (int k = 0; k < 4; k++) KRT_HIP(hipEventElapsedTime(&times.ms[k], ev[k], ev[k + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
}

}  // namespace krt

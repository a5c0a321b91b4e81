// Path tracer attachment: the estimator of pt_pixel (krt_pt.h,
// amadeus_path_tracer.cu:73-228) in wavefront form, held equal to pt_pixel
// bit for bit by the serving tests.
#include <hip/hip_runtime.h>

#include "krt_cache.h"
#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_shadow.h"
#include "krt_traverse.h"

namespace krt {

namespace {

// The path is split at every ray.  Per level L,
//   walk_closest_list     the level's rays (slot = pixel) -> hits
//   pt_shade_kernel       surface, material, the light sample of Ld and its
//                         shadow ray, the BSDF sample -> the next level's ray,
//                         the level's history {Le, has_pdf}, {T, Ld flag},
//                         {Ld contribution unblocked, Ld pdf}; level 0 writes
//                         the normal / albedo / position AOVs
//   walk_shadow_list      -> occlusion byte of (L, pixel)
// and pt_compose_kernel folds the history from the deepest level up and
// accumulates the colour (and the FP64 sums).  Every number comes from the
// expressions of pt_pixel in the same order, the random sequence included,
// so the two agree bit for bit; the walks run in the register-lean list
// kernels at up to eight waves per SIMD.
struct PtWave {
    RayList cl, sl;   // closest and shadow rays of the current level
    float4 *hits;     // per pixel: {t, b1, b2, prim bits}
    int *hit_inst;    // per pixel: instance (two-level scenes)
    float4 *state;    // per pixel: {seed, levels with history (int bits)}
    float4 *hist;     // per level, per pixel: 3 float4
    uint8_t *occ;     // per level, per pixel
    float4 *lec;      // radiance cache: the emission added exactly at the cache vertex
};

constexpr int kWaveBlock = 256;
// state.w: levels with history in the low 16 bits, bit 16 set when the path
// recorded a radiance-cache query
constexpr int kQueried = 1 << 16;
// state.w bit: the path's last level is a miss that took the environment's radiance
constexpr int kEnvMiss = 1 << 17;

__global__ __launch_bounds__(kWaveBlock) void pt_camera_kernel(LaunchInfo li, PtWave B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t px = 0, py = 0, pix = 0;
    const bool live = item < tiles.items && tiles.pixel(item, px, py);
    f3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    if (live) {
        pix = py * li.width + px;
        f3 seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
        make_ray(li, px, py, seed, li.jitter != 0, o, d);
        B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float(0));
    }
    emit_ray(B.cl, &nray, live, o, d, kTraceTmax, pix);
    publish_rays(B.cl, &nray);
}

template <bool INST>
__global__ __launch_bounds__(kWaveBlock) void pt_shade_kernel(LaunchInfo li, SceneView sv, PtWave B, int level,
                                                              CacheView cv) {
    __shared__ uint32_t ncl, nsl;
    if (threadIdx.x == 0) ncl = nsl = 0;
    __syncthreads();
    const uint32_t npix = li.width * li.height;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const float count = (float)li.samples;
    const bool acc = li.accumulate != 0;
    const EnvMap env = env_of(li);
    const bool live = k < B.cl.prefix[B.cl.nseg];
    bool shadow = false, next = false;
    f3 so = mk(0, 0, 0), sd = mk(0, 0, 0), no = mk(0, 0, 0), nd = mk(0, 0, 0);
    float stmax = 0.0f;
    uint32_t pix = 0;
    if (live) {
        const float4 rb = B.cl.dense[2 * (size_t)k + 1];
        const f3 d = mk(rb.x, rb.y, rb.z);
        pix = __float_as_uint(rb.w);
        const float4 st = B.state[pix];
        f3 seed = mk(st.x, st.y, st.z);
        const float4 hh = B.hits[pix];
        Hit h;
        h.t = hh.x;
        h.b1 = hh.y;
        h.b2 = hh.z;
        h.prim = __float_as_int(hh.w);
        h.rank = 0;
        h.inst = INST ? B.hit_inst[pix] : -1;
        Surface sf;
        const bool surf = h.prim >= 0 && load_surface<INST>(sv, h, d, sf);  // a miss or a NaN normal ends the path
        int queried = __float_as_int(st.w) & kQueried;
        if (h.prim < 0 && env.tex) {  // __miss__ (amadeus_path_tracer.cu:204-220): value = the environment's
            const f3 Le = env_radiance(env, d, true);
            float4 *hl = B.hist + 3 * (size_t)level * npix;
            hl[pix] = make_float4(Le.x, Le.y, Le.z, 0.0f);
            hl[npix + pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            hl[2 * (size_t)npix + pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            B.state[pix] = make_float4(st.x, st.y, st.z, __int_as_float((level + 1) | queried | kEnvMiss));
        }
        bool cache_stop = false;
        if (surf && cv.enabled && level == cv.depth) {
            // the cache vertex (krt_cache.hip): record the query; the path stops
            // here unless the pixel is a training path
            const Material mat = surface_material(sv, sf);
            uint32_t slot = 0;
            const bool train = cv.training(pix, slot);
            put_query(train ? cv.slot_raw + (size_t)kMlpRaw * slot : cv.query + (size_t)kMlpRaw * pix, sf.x, -d, sf.n,
                      mat);
            const f3 le = cv.learn_emission ? mk(0, 0, 0) : mat.emission;
            B.lec[pix] = make_float4(le.x, le.y, le.z, 0.0f);
            queried = kQueried;
            B.state[pix] = make_float4(st.x, st.y, st.z, __int_as_float(__float_as_int(st.w) | kQueried));
            cache_stop = !train;
        }
        if (surf && !cache_stop) {
            const Material mat = surface_material(sv, sf);
            const f3 n = sf.n;
            const f3 x = offset_point(sf.x, n, mat.type);
            const SurfaceHit sh{mat, sf.entering, n, -d, x};
            // Ld (krt_shade.h) with its shadow ray deferred
            f3 contr = mk(0, 0, 0);
            float lpdf = 0.0f;
            LdSample ls;
            const bool ld = ld_sample(sv, env, sh, seed, ls);
            if (ld) {
                contr = ls.contr;
                lpdf = ls.pdf;
                so = sh.x;
                sd = ls.wi;
                stmax = ls.tmax;
                shadow = true;
            }
            int out;
            f3 wi;
            float pdf;
            f3 f = eval_bsdf(sh, wi, pdf, out, seed);
            f3 T = f * fabsf(dot(wi, n)) / pdf;
            bool has_pdf = pdf > 0.0f;
            float4 *hl = B.hist + 3 * (size_t)level * npix;
            hl[pix] = make_float4(mat.emission.x, mat.emission.y, mat.emission.z, has_pdf ? 1.0f : 0.0f);
            hl[npix + pix] = make_float4(T.x, T.y, T.z, ld ? 1.0f : 0.0f);
            hl[2 * (size_t)npix + pix] = make_float4(contr.x, contr.y, contr.z, lpdf);
            if (level == 0) {
                li.normal[pix] = accumulate4(li.normal[pix], make_float4(n.x, n.y, n.z, 0.0f), acc, count);
                li.albedo[pix] =
                    accumulate4(li.albedo[pix], make_float4(mat.diffuse.x, mat.diffuse.y, mat.diffuse.z, 0.0f), acc,
                                count);
                li.position[pix] = make_float4(x.x, x.y, x.z, 0.0f);
            }
            B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float((level + 1) | queried));
            if (has_pdf && level < li.max_depth) {
                no = x;
                nd = wi;
                next = true;
            }
        }
    }
    emit_ray(B.sl, &nsl, shadow, so, sd, stmax, pix);
    emit_ray(B.cl, &ncl, next, no, nd, kTraceTmax, pix);
    __syncthreads();
    if (threadIdx.x == 0) {
        B.sl.counts[xcd_block()] = nsl;
        B.cl.counts[xcd_block()] = ncl;
    }
}

// Le + Ld of level k of pixel pix, the shadow ray's answer applied.
KRT_DEV f3 level_direct(const PtWave &B, uint32_t npix, uint32_t pix, int k, float4 &e, float4 &t) {
    const float4 *hl = B.hist + 3 * (size_t)k * npix;
    e = hl[pix];
    t = hl[npix + pix];
    const float4 c = hl[2 * (size_t)npix + pix];
    f3 Ldv = mk(0.0f, 0.0f, 0.0f);
    if (t.w != 0.0f) {
        const f3 contr = B.occ[(size_t)k * npix + pix] ? mk(0.0f, 0.0f, 0.0f) : mk(c.x, c.y, c.z);
        Ldv = contr / c.w;
    }
    return mk(e.x, e.y, e.z) + Ldv;
}

// With the radiance cache (krt_cache.hip): a path that stopped at the cache
// vertex leaves value_0 = P + Theta * (Le + cache) as prefix P (Le folded in)
// and Theta for the composite after inference; any other path leaves its
// value, and a training path its target (the value at the cache vertex less
// that vertex's emission).  The AOVs are written as without the cache.
__global__ __launch_bounds__(kWaveBlock) void pt_cache_compose_kernel(LaunchInfo li, PtWave B, CacheView cv,
                                                                      float4 *prefix) {
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const int sw = __float_as_int(B.state[pix].w);
    const int nlev = sw & (kQueried - 1);
    const bool queried = (sw & kQueried) != 0;
    uint32_t slot = 0;
    const bool train = cv.training(pix, slot);
    f3 le_c = mk(0, 0, 0);
    if (queried) {
        const float4 l = B.lec[pix];
        le_c = mk(l.x, l.y, l.z);
    }
    float4 pre, th = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (queried && !train) {
        f3 P = mk(0, 0, 0), Th = mk(1.0f, 1.0f, 1.0f);
        for (int k = nlev - 1; k >= 0; k--) {
            float4 e, t;
            const f3 dk = level_direct(B, npix, pix, k, e, t);
            const f3 T = mk(t.x, t.y, t.z);
            P = dk + T * P;
            Th = T * Th;
        }
        P = P + Th * le_c;
        pre = make_float4(P.x, P.y, P.z, 0.0f);
        th = make_float4(Th.x, Th.y, Th.z, 1.0f);
    } else {
        f3 value = mk(0, 0, 0), target = mk(0, 0, 0);
        for (int k = nlev - 1; k >= 0; k--) {
            float4 e, t;
            f3 v = level_direct(B, npix, pix, k, e, t);
            if (e.w != 0.0f) v = v + mk(t.x, t.y, t.z) * value;
            value = v;
            if (k == cv.depth) target = value - le_c;  // scattered radiance only
        }
        pre = make_float4(value.x, value.y, value.z, 0.0f);
        if (train) {
            float *tg = cv.slot_tgt + 3 * (size_t)slot;
            tg[0] = target.x;
            tg[1] = target.y;
            tg[2] = target.z;
            cv.slot_ok[slot] = (queried && nlev > cv.depth && finite3(target)) ? 1u : 0u;
        }
    }
    prefix[pix] = pre;
    cv.theta[pix] = th;
    if (nlev == 0 || (nlev == 1 && (sw & kEnvMiss))) {  // the primary ray missed: the AOVs take zeros
        const bool acc = li.accumulate != 0;
        const float count = (float)li.samples;
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        li.normal[pix] = accumulate4(li.normal[pix], zero, acc, count);
        li.albedo[pix] = accumulate4(li.albedo[pix], zero, acc, count);
        li.position[pix] = zero;
    }
}

__global__ __launch_bounds__(kWaveBlock) void pt_compose_kernel(LaunchInfo li, PtWave B, double *acc64) {
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const int sw = __float_as_int(B.state[pix].w);
    const int nlev = sw & (kQueried - 1);
    f3 value = mk(0, 0, 0);
    for (int k = nlev - 1; k >= 0; k--) {
        float4 e, t;
        f3 v = level_direct(B, npix, pix, k, e, t);
        if (e.w != 0.0f) v = v + mk(t.x, t.y, t.z) * value;
        value = v;
    }
    float4 col = make_float4(value.x, value.y, value.z, 1.0f);
    if (isnan3(value)) col = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    else if (acc64) {  // converged-reference accumulator (FP64 only here)
        double *sum = acc64 + 3 * (size_t)pix;
        double *sq = acc64 + 3 * ((size_t)npix + pix);  // per-pixel sample variance
        const double v[3] = {(double)value.x, (double)value.y, (double)value.z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sum[c] += v[c];
            sq[c] += v[c] * v[c];
        }
    }
    const bool acc = li.accumulate != 0;
    const float count = (float)li.samples;
    li.color[pix] = accumulate4(li.color[pix], col, acc, count);
    if (nlev == 0 || (nlev == 1 && (sw & kEnvMiss))) {  // the primary ray missed: the AOVs take zeros, as pt_pixel's
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        li.normal[pix] = accumulate4(li.normal[pix], zero, acc, count);
        li.albedo[pix] = accumulate4(li.albedo[pix], zero, acc, count);
        li.position[pix] = zero;
    }
}

}  // namespace

// Workgroups of the wavefront passes: whole 8x8 tiles, 256 lanes a group.
static uint32_t wave_blocks(uint32_t w, uint32_t h) {
    const uint32_t items = ((w + 7u) / 8u) * ((h + 7u) / 8u) * 64u;
    return (items + kWaveBlock - 1) / kWaveBlock;
}

// The size of each wavefront-scratch buffer for a frame of w x h pixels:
// hits, instance ids, path state, per-level history (3 float4) and occlusion
// bytes, cache-vertex emission, and per ray list (closest, shadow) the rays,
// segment counts, offsets and the packed copy.  `b` is filled in the order of
// the buffers of WaveScratch; zero means the buffer is not needed.
static void wave_sizes(uint32_t w, uint32_t h, int max_depth, bool instanced, bool cache, size_t b[14]) {
    const size_t npix = (size_t)w * h;
    const uint32_t gb = wave_blocks(w, h);
    const size_t levels = (size_t)max_depth + 1;
    const size_t cap = (size_t)gb * kWaveBlock;
    b[0] = sizeof(float4) * npix;
    b[1] = instanced ? sizeof(int) * npix : 0;
    b[2] = sizeof(float4) * npix;
    b[3] = sizeof(float4) * 3 * levels * npix;
    b[4] = levels * npix;
    b[5] = cache ? sizeof(float4) * npix : 0;
    for (int l = 0; l < 2; l++) {
        b[6 + 4 * l] = sizeof(float4) * 2 * cap;
        b[7 + 4 * l] = sizeof(uint32_t) * gb;
        b[8 + 4 * l] = sizeof(uint32_t) * (gb + 1);
        b[9 + 4 * l] = sizeof(float4) * 2 * cap;
    }
}

static void wave_bufs(WaveScratch &ws, DevBuf *b[14]) {
    b[0] = &ws.hits;
    b[1] = &ws.inst;
    b[2] = &ws.state;
    b[3] = &ws.hist;
    b[4] = &ws.occ;
    b[5] = &ws.lec;
    for (int l = 0; l < 2; l++) {
        b[6 + 4 * l] = &ws.rays[l];
        b[7 + 4 * l] = &ws.counts[l];
        b[8 + 4 * l] = &ws.prefix[l];
        b[9 + 4 * l] = &ws.dense[l];
    }
}

size_t wave_scratch_grow(WaveScratch &ws, uint32_t w, uint32_t h, int max_depth, bool instanced, bool cache,
                         bool dry_run) {
    size_t need[14];
    DevBuf *bufs[14];
    wave_sizes(w, h, max_depth, instanced, cache, need);
    wave_bufs(ws, bufs);
    size_t grown = 0;
    for (int i = 0; i < 14; i++)
        if (need[i] && bufs[i]->bytes < need[i]) {
            grown += need[i] - bufs[i]->bytes;
            if (!dry_run) bufs[i]->alloc(need[i]);
        }
    return grown;
}

size_t wave_scratch_bytes(const WaveScratch &ws) {
    DevBuf *bufs[14];
    wave_bufs(const_cast<WaveScratch &>(ws), bufs);
    size_t b = 0;
    for (DevBuf *d : bufs) b += d->bytes;
    return b;
}

void pt_wave(Context &ctx, Scene &s, const LaunchInfo &li, unsigned long long *ctr, double *acc64,
             const CacheView *cv, float4 *prefix, hipEvent_t queried) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(li.max_depth >= 0 && li.max_depth <= 64, KRT_ERR_INVALID, "max_depth must be in [0, 64]");
    const uint32_t npix = li.width * li.height;
    const CacheView cview = cv ? *cv : CacheView{};
    const SceneView sv = make_view(s);
    WaveScratch &ws = ctx.wave;
    const uint32_t gb = wave_blocks(li.width, li.height);
    wave_scratch_grow(ws, li.width, li.height, li.max_depth, s.instanced, cview.enabled);
    spill_buffer(ctx, s, 2 * kMaxResidentBlocks * kTraceBlock, kLdsDepthLean);
    int *spill = ctx.spill.as<int>();
    PtWave B;
    B.cl = RayList{ws.rays[0].as<float4>(), ws.counts[0].as<uint32_t>(), (uint32_t)kWaveBlock, gb,
                   ws.prefix[0].as<uint32_t>(), ws.dense[0].as<float4>()};
    B.sl = RayList{ws.rays[1].as<float4>(), ws.counts[1].as<uint32_t>(), (uint32_t)kWaveBlock, gb,
                   ws.prefix[1].as<uint32_t>(), ws.dense[1].as<float4>()};
    B.hits = ws.hits.as<float4>();
    B.hit_inst = s.instanced ? ws.inst.as<int>() : nullptr;
    B.state = ws.state.as<float4>();
    B.hist = ws.hist.as<float4>();
    B.occ = ws.occ.as<uint8_t>();
    B.lec = cview.enabled ? ws.lec.as<float4>() : nullptr;
    pt_camera_kernel<<<gb, kWaveBlock, 0, ctx.stream>>>(li, B);
    KRT_HIP(hipGetLastError());
    walk_closest_list(ctx, s.instanced, sv, B.cl, B.hits, B.hit_inst, spill, ctr, 0);
    // Past the cache vertex only the training paths (at most cv.M) go on, so
    // those levels' shading grids, and the segments of the lists they emit,
    // shrink to the training paths' count; with no training paths they end.
    const uint32_t train_blocks = (cview.M + kWaveBlock - 1) / kWaveBlock;
    uint32_t in_seg = gb;  // segments of the closest list the next shading pass reads
    for (int level = 0; level <= li.max_depth; level++) {
        uint32_t grid = gb;
        if (cview.enabled && level > cview.depth) {
            if (train_blocks == 0) break;
            grid = std::min(gb, train_blocks);
        }
        PtWave Bl = B;
        Bl.cl.nseg = in_seg;
        if (s.instanced) pt_shade_kernel<true><<<grid, kWaveBlock, 0, ctx.stream>>>(li, sv, Bl, level, cview);
        else pt_shade_kernel<false><<<grid, kWaveBlock, 0, ctx.stream>>>(li, sv, Bl, level, cview);
        KRT_HIP(hipGetLastError());
        if (queried && cview.enabled && level == cview.depth) {
            KRT_HIP(hipEventRecord(queried, ctx.stream));
            queried = nullptr;
        }
        Bl.cl.nseg = Bl.sl.nseg = in_seg = grid;
        // the level's shadow rays and the next level's closest rays
        if (level < li.max_depth)
            walk_pair(ctx, s.instanced, sv, Bl.cl, Bl.hits, Bl.hit_inst, Bl.sl, Bl.occ + (size_t)level * npix, spill,
                      ctr, 0);
        else
            walk_shadow_list(ctx, s.instanced, sv, Bl.sl, Bl.occ + (size_t)level * npix, spill, ctr, 0);
    }
    if (cview.enabled) pt_cache_compose_kernel<<<gb, kWaveBlock, 0, ctx.stream>>>(li, B, cview, prefix);
    else pt_compose_kernel<<<gb, kWaveBlock, 0, ctx.stream>>>(li, B, acc64);
    KRT_HIP(hipGetLastError());
    if (queried) KRT_HIP(hipEventRecord(queried, ctx.stream));  // the cache vertex lies beyond the path's depth
}

void pt_render(Context &ctx, Scene &s, const LaunchInfo &li, unsigned long long *ctr, PassTimes &times,
               double *acc64) {
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, ctx.stream));
    pt_wave(ctx, s, li, ctr, acc64, nullptr, nullptr);
    KRT_HIP(hipEventRecord(e1, ctx.stream));
    KRT_HIP(hipEventSynchronize(e1));
    KRT_HIP(hipEventElapsedTime(&times.ms[0], e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace krt

// ReSTIR DI attachment: the three per-pixel passes of
// source/optix/amadeus_restir.cu with the host sequence of
// include/amadeus/restir.cuh:223-304.  Each pass is split at its rays into
// producer / consumer kernels around the ray-list walks of krt_shadow.h (the
// initial pass with the radiance cache is one fused kernel with path
// regeneration).
//
//   initial  (__raygen__initial + __closesthit__initial, :82-144, :437-559):
//            primary hit -> RIS over `candidates` light samples (target = sum
//            of unshadowed RGB), one visibility ray, G-buffer material id; the
//            rest of the path adds Le + Ld per bounce (intermediate buffer).
//   temporal (__raygen__temporal, :213-291): merge with the previous frame's
//            reservoir (M capped), occluded targets, resample.
//   spatial  (__raygen__spatial, :317-434): `neighbours` reservoirs in a disk
//            of `radius` px, 1/Z bias correction, shade, accumulate.
//
// Reservoirs (include/amadeus/reservoir.cuh:19-76) are stored as four float4
// planes per buffer: {Le, type}, {normal, w}, {point, W}, {seed, M}.
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

constexpr int kMaxNeighbours = 8;

struct RSample {
    f3 Le, normal, point;
    int type;
};

struct Resv {
    RSample data;
    float w, W;
    int M;
    f3 seed;
    KRT_DEV bool update(const RSample &s, float weight) {
        w += weight;
        float eta = rand_uniform(seed);
        bool selected = (eta * w) < weight;
        if (selected) data = s;
        M++;
        return selected;
    }
    KRT_DEV void resample(float target) {
        float d = target * (float)M;
        W = (d > 0.0f) ? w / d : 0.0f;
    }
};

// A frame's reservoirs: one 64 B record of four float4 per pixel, so a
// neighbour's reservoir is one cache line to fetch.
struct Planes {
    float4 *p;
    uint32_t n;
    KRT_DEV float4 &part(uint32_t k, uint32_t i) const { return p[4 * (size_t)i + k]; }
    KRT_DEV Resv load(uint32_t i) const {
        float4 a = part(0, i), b = part(1, i), c = part(2, i), d = part(3, i);
        Resv r;
        r.data.Le = mk(a.x, a.y, a.z);
        r.data.type = __float_as_int(a.w);
        r.data.normal = mk(b.x, b.y, b.z);
        r.w = b.w;
        r.data.point = mk(c.x, c.y, c.z);
        r.W = c.w;
        r.seed = mk(d.x, d.y, d.z);
        r.M = __float_as_int(d.w);
        return r;
    }
    KRT_DEV void store(uint32_t i, const Resv &r) const {
        part(0, i) = make_float4(r.data.Le.x, r.data.Le.y, r.data.Le.z, __int_as_float(r.data.type));
        part(1, i) = make_float4(r.data.normal.x, r.data.normal.y, r.data.normal.z, r.w);
        part(2, i) = make_float4(r.data.point.x, r.data.point.y, r.data.point.z, r.W);
        part(3, i) = make_float4(r.seed.x, r.seed.y, r.seed.z, __int_as_float(r.M));
    }
    // Reservoir::reset (reservoir.cuh:64-69): w = M = W = 0, data and seed kept
    KRT_DEV void reset(uint32_t i) const {
        part(1, i).w = 0.0f;
        part(2, i).w = 0.0f;
        part(3, i).w = __int_as_float(0);
    }
    KRT_DEV void store_M(uint32_t i, int M) const { part(3, i).w = __int_as_float(M); }
    KRT_DEV int M(uint32_t i) const { return __float_as_int(part(3, i).w); }
};

// direct_unoccluded (amadeus_common.cuh:380-397)
KRT_DEV f3 direct_li(const SurfaceHit &sh, f3 Le, f3 normal, int type, f3 D, float d) {
    f3 rho = brdf(sh, D, eDiffuse);
    float geometric = fabsf(dot(sh.n, D));
    if (type != 2) {
        float ldot = fabsf(dot(normal, D));
        geometric *= ldot / (d * d);
    }
    return rho * Le * geometric;
}

// direct_occluded (:399-421)
template <bool WIDE = true, bool INST = false>
KRT_DEV f3 direct_occ(const SceneView &sv, const SurfaceHit &sh, const RSample &s, f3 D, float d,
                      const TraceStack &stk, TravStats &ts, RayCount &rc, bool *occ_out = nullptr) {
    rc.shadow++;
    bool occ = occluded<WIDE, INST>(sv, sh.x, D, d - kShadowEps, stk, ts);
    if (occ_out) *occ_out = occ;
    if (occ) return mk(0.0f, 0.0f, 0.0f);
    return direct_li(sh, s.Le, s.normal, s.type, D, d);
}

// target / occluded_target (amadeus_restir.cu:148-178)
KRT_DEV float target(const RSample &s, const SurfaceHit &hit) {
    f3 D = s.point - hit.x;
    float d = length(D);
    f3 Li = direct_li(hit, s.Le, s.normal, s.type, D / d, d);
    return Li.x + Li.y + Li.z;
}

// A resampling ray whose answer cannot change anything: the candidate's
// weight is target * W * M, and with W = 0 and a finite unshadowed target it
// is 0 whether the ray is blocked or not (a zero-weight candidate is never
// selected).  Producers skip such rays; consumers take them as unblocked.
KRT_DEV bool moot_ray(float W, f3 unshadowed) {
    return W == 0.0f && isfinite(unshadowed.x + unshadowed.y + unshadowed.z);
}

template <bool INST>
KRT_DEV float occluded_target(const SceneView &sv, const RSample &s, const SurfaceHit &hit, const TraceStack &stk,
                              TravStats &ts, RayCount &rc) {
    f3 D = s.point - hit.x;
    float d = length(D);
    f3 Li = direct_occ<true, INST>(sv, hit, s, D / d, d, stk, ts, rc);
    return Li.x + Li.y + Li.z;
}

struct RestirArgs {
    LaunchInfo li;
    SceneView sv;
    Planes cur, prev;  // prev: temporal source / spatial destination
    uint32_t *gmat;
    float2 *guv;  // textured scenes: the G-buffer's texture coordinates
    float4 *intermediate, *auxiliary;
    int *spill;
    unsigned long long *ctr;
    int candidates, m_cap, neighbours, radius;
    CacheView cv;  // radiance cache for the indirect path (cv.enabled = 0: off)
};

// The G-buffer's material of a pixel (parameters.materials_buffer,
// amadeus_restir.cu:231): its id and, for a textured material, its texture
// coordinates.
KRT_DEV Material gbuffer_material(const RestirArgs &A, uint32_t i) {
    const uint32_t m = A.gmat[i];
    Material M = load_material(A.sv.mats, m);
    const uint4 t = material_tex(A.sv.mats, m);
    if (t.w) {
        const float2 uv = A.guv[i];
        apply_textures(A.sv.tex, t, uv.x, uv.y, M);
    }
    return M;
}

// G-buffer surface of a pixel (amadeus_restir.cu:230-240, :323-336); the
// normal is the accumulated normal AOV, as in the reference.
KRT_DEV SurfaceHit gbuffer_hit(const RestirArgs &A, uint32_t i) {
    SurfaceHit h;
    h.mat = gbuffer_material(A, i);
    h.entering = false;
    float4 n = A.li.normal[i], x = A.li.position[i];
    h.n = mk(n.x, n.y, n.z);
    h.x = mk(x.x, x.y, x.z);
    h.wo = normalize(ld3(A.li.center) - h.x);
    return h;
}

// reproject (amadeus_restir.cu:182-210); pixel indices clamped to the image.
KRT_DEV int reproject(const LaunchInfo &li, int index, f3 pos) {
    if (li.samples > 0) return index;
    const float *m = li.prev_pv;
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (m[0 + k] * pos.x + m[4 + k] * pos.y) + (m[8 + k] * pos.z + m[12 + k] * 1.0f);
    float u = v[0] / v[3], vv = v[1] / v[3];
    if (!(u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f)) return -1;
    u = (u + 1.0f) * 0.5f;
    vv = (vv + 1.0f) * 0.5f;
    int ix = round_half_up(u * (float)li.width);
    int iy = round_half_up(vv * (float)li.height);
    ix = min(max(ix, 0), (int)li.width - 1);
    iy = min(max(iy, 0), (int)li.height - 1);
    return iy * (int)li.width + ix;
}

// The spatial pass (__raygen__spatial, amadeus_restir.cu:317-434) split around
// its shadow rays.  Per pixel the reference traces one ray towards every
// non-empty neighbour's sample (the resampling target), then one from every
// non-empty neighbour's surface towards the merged sample (1/Z), then the
// shading ray.  Here:
//   spatial_rays_kernel     neighbour picks -> ray list (slot i = neighbour i)
//   shadow_list_kernel      any-hit walks of the list -> occ[slot][pixel]
//   spatial_merge_kernel    resampling with those visibilities; 1/Z rays
//                           (slot i) and the shading ray (slot S) -> ray list
//   shadow_list_kernel
//   spatial_resolve_kernel  Z, W, shading, accumulation
// Every ray has the origin, direction and extent of the reference's, walked
// by the same traversal, so each visibility is the one the fused loop saw;
// the walks run in a register-lean kernel at eight waves per SIMD instead of
// inside the shading code's 128-VGPR frame.
//
// The neighbour picks are replayed from the pixel's auxiliary seed in each
// kernel: a pick draws three numbers and every non-empty neighbour's
// reservoir update one more, whatever its visibility, so the seed sequence is
// known before any ray is traced.

struct SpatialBufs {
    RayList list;
    uint8_t *occ;    // [slot * npix + pixel]: 1 when the slot's ray was blocked
    int32_t *zm;     // [i * npix + pixel]: neighbour i's M when its unoccluded 1/Z target is > 0, else 0
    uint8_t *flags;  // per pixel: 1 sample selected from a neighbour whose ray is known, 2 that ray was blocked
};

constexpr int kSpatialBlock = 256;

// The initial pass (__raygen__initial + __closesthit__initial,
// amadeus_restir.cu:82-144, :437-559) split at every ray.  Per level L of
// the path,
//   walk_closest_list / walk_pair   the level's rays (slot = pixel) -> hits
//   init_shade_kernel<L>   surface, material; L = 0: RIS over the
//                          candidates, the reservoir with W for an unblocked
//                          sample, the sample's visibility ray; L >= 1: the
//                          light sample of Ld and its shadow ray; the BSDF
//                          sample -> the next level's ray; per-level history.
//                          With the radiance cache, the level `cache_depth`
//                          hit records the MLP query and ends the path
//                          (training paths go on to full depth).
//   walk_shadow_list / walk_pair    -> occlusion byte of (L, pixel)
// and init_compose_kernel folds the history from the deepest level up
// (value_k = direct_k + T_k value_k+1, or P + Theta (Le + cache) with the
// cache), zeroes W of reservoirs whose sample is blocked, and writes the
// intermediate and auxiliary buffers.  Every number is computed by the
// expressions of the reference's programs in the same order, the random
// sequence included (the oracle holds it bit for bit); the walks run in the
// register-lean list kernels.
struct InitWave {
    RayList cl;      // closest rays of the current level
    RayList sl;      // shadow rays of the current level
    float4 *hits;    // per pixel: {t, b1, b2, prim bits} of the current level's ray
    int *hit_inst;   // per pixel: instance of that hit (two-level scenes)
    float4 *state;   // per pixel: {seed, levels with history (int bits)}
    float4 *hist;    // per level, per pixel: {emission, 0}, {T, Ld flag}, {Ld contribution unblocked, Ld pdf}
    uint8_t *occ;    // per level, per pixel
    float4 *lec;     // radiance cache: per pixel, the emission added exactly at the cache vertex
};

// state.w of the wavefront initial pass: levels with history in the low 16
// bits, bit 16 set when the path recorded a radiance-cache query
constexpr int kQueried = 1 << 16;
// state.w bit: the path's last level is a miss that took the environment's radiance
constexpr int kEnvMiss = 1 << 17;

__global__ __launch_bounds__(kSpatialBlock) void init_camera_kernel(RestirArgs A, InitWave B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t px = 0, py = 0, pix = 0;
    const bool live = item < tiles.items && tiles.pixel(item, px, py);
    f3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    if (live) {
        pix = py * li.width + px;
        f3 seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
        make_ray(li, px, py, seed, false, o, d);
        B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float(0));
    }
    emit_ray(B.cl, &nray, live, o, d, kTraceTmax, pix);
    publish_rays(B.cl, &nray);
}

template <bool INST>
__global__ __launch_bounds__(kSpatialBlock) void init_shade_kernel(RestirArgs A, InitWave B, int level) {
    __shared__ uint32_t ncl, nsl;
    if (threadIdx.x == 0) ncl = nsl = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const SceneView &sv = A.sv;
    const uint32_t npix = li.width * li.height;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const float count = (float)li.samples;
    const bool acc = li.accumulate != 0;
    const EnvMap env = env_of(li);
    const unsigned total = sv.nl + (env.tex ? 1u : 0u);
    const bool live = k < B.cl.prefix[B.cl.nseg];
    bool shadow = false, next = false;
    f3 so = mk(0, 0, 0), sd = mk(0, 0, 0), no = mk(0, 0, 0), nd = mk(0, 0, 0);
    float stmax = 0.0f;
    uint32_t pix = 0;
    if (live) {
        const float4 ra = B.cl.dense[2 * (size_t)k], rb = B.cl.dense[2 * (size_t)k + 1];
        (void)ra;
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
        const CacheView &cv = A.cv;
        const bool surf = h.prim >= 0 && load_surface<INST>(sv, h, d, sf);
        int queried = __float_as_int(st.w) & kQueried;
        if (h.prim < 0 && env.tex) {  // __miss__ (amadeus_restir.cu:566-577): the environment's radiance
            const f3 Le = env_radiance(env, d, false);
            float4 *hl = B.hist + 3 * (size_t)level * npix;
            hl[pix] = make_float4(Le.x, Le.y, Le.z, 0.0f);
            hl[npix + pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            hl[2 * (size_t)npix + pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            B.state[pix] = make_float4(st.x, st.y, st.z, __int_as_float((level + 1) | queried | kEnvMiss));
        }
        bool cache_stop = false;
        if (surf && cv.enabled && level == cv.depth) {
            // the cache vertex: record the query; the path stops here unless
            // the pixel is a training path
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
            f3 emission = mk(0, 0, 0), contr = mk(0, 0, 0);
            float lpdf = 0.0f;
            bool ld = false;
            if (level == 0) {
                Resv res = A.cur.load(pix);
                res.w = 0.0f;
                res.M = 0;
                res.W = 0.0f;
                for (int i = 0; i < A.candidates; i++) {
                    // sample_direct (amadeus_common.cuh:301-377): triangle lights and the environment
                    RSample smp;
                    smp.Le = smp.normal = smp.point = mk(0, 0, 0);
                    smp.type = 0;
                    float pdf = 0.0f;
                    if (total > 0) {
                        unsigned li_ = pick_light(seed, total);
                        if (li_ < sv.nl) {
                            TriLight L = load_light(sv.lights, li_);
                            smp.point = sample_area_light(L, seed);
                            smp.Le = L.intensity;
                            smp.normal = L.normal;
                            smp.type = 1;
                            pdf = 1.0f / (L.area * (float)total);
                        } else {  // a direction toward the environment, WORLD_RADIUS away
                            const f3 wi = env_sample_dir(seed);
                            smp.point = sh.x + wi * kEnvDistance;
                            smp.Le = env_radiance(env, wi, false);
                            smp.normal = -wi;
                            smp.type = 2;
                            pdf = 1.0f / (kFourPiF * (float)total);
                        }
                    } else {
                        rand_uniform(seed);
                    }
                    f3 D = smp.point - sh.x;
                    float dd = length(D);
                    f3 Li = direct_li(sh, smp.Le, smp.normal, smp.type, D / dd, dd);
                    float tgt = Li.x + Li.y + Li.z;
                    res.update(smp, (pdf > 0.0f) ? tgt / pdf : 0.0f);
                }
                const RSample smp = res.data;
                const f3 D = smp.point - sh.x;
                const float dd = length(D);
                const f3 Li = direct_li(sh, smp.Le, smp.normal, smp.type, D / dd, dd);
                res.resample(Li.x + Li.y + Li.z);  // W for an unblocked sample; the compose kernel applies the ray
                A.cur.store(pix, res);
                A.gmat[pix] = sf.mat_id;
                if (sv.uv) A.guv[pix] = make_float2(sf.u, sf.v);
                so = sh.x;
                sd = D / dd;
                stmax = dd - kShadowEps;
                shadow = true;
            } else {
                // Ld (krt_shade.h) with its shadow ray deferred
                emission = mat.emission;
                LdSample ls;
                ld = ld_sample(sv, env, sh, seed, ls);
                if (ld) {
                    contr = ls.contr;
                    lpdf = ls.pdf;
                    so = sh.x;
                    sd = ls.wi;
                    stmax = ls.tmax;
                    shadow = true;
                }
            }
            int out;
            f3 wi;
            float pdf;
            f3 f = eval_bsdf(sh, wi, pdf, out, seed);
            bool has_pdf = pdf > 0.0f;
            f3 T = has_pdf ? f * fabsf(dot(wi, n)) / pdf : mk(0.0f, 0.0f, 0.0f);
            float4 *hl = B.hist + 3 * (size_t)level * npix;
            hl[pix] = make_float4(emission.x, emission.y, emission.z, 0.0f);
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

__global__ __launch_bounds__(kSpatialBlock) void init_compose_kernel(RestirArgs A, InitWave B) {
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const float4 st = B.state[pix];
    const int sw = __float_as_int(st.w);
    const int nlev = sw & (kQueried - 1);
    const bool queried = (sw & kQueried) != 0;
    const CacheView &cv = A.cv;
    uint32_t slot = 0;
    const bool train = cv.enabled && cv.training(pix, slot);
    f3 le_c = mk(0, 0, 0);
    if (queried) {
        const float4 l = B.lec[pix];
        le_c = mk(l.x, l.y, l.z);
    }
    f3 value = mk(0, 0, 0), Th = mk(1.0f, 1.0f, 1.0f), target = mk(0, 0, 0);
    for (int k = nlev - 1; k >= 0; k--) {
        const float4 *hl = B.hist + 3 * (size_t)k * npix;
        const float4 e = hl[pix], t = hl[npix + pix], c = hl[2 * (size_t)npix + pix];
        f3 Ldv = mk(0.0f, 0.0f, 0.0f);
        if (t.w != 0.0f) {
            const f3 contr = B.occ[(size_t)k * npix + pix] ? mk(0.0f, 0.0f, 0.0f) : mk(c.x, c.y, c.z);
            Ldv = contr / c.w;
        }
        const f3 dk = mk(e.x, e.y, e.z) + Ldv;
        const f3 T = mk(t.x, t.y, t.z);
        value = dk + T * value;
        Th = T * Th;
        if (cv.enabled && k == cv.depth) target = value - le_c;
    }
    if (queried && !train) {
        // value = P + Theta * (Le_c + cache); the cache term is folded in after inference
        value = value + Th * le_c;
        cv.theta[pix] = make_float4(Th.x, Th.y, Th.z, 1.0f);
    } else if (cv.enabled) {
        cv.theta[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (train) {
            float *tg = cv.slot_tgt + 3 * (size_t)slot;
            tg[0] = target.x;
            tg[1] = target.y;
            tg[2] = target.z;
            cv.slot_ok[slot] = (queried && nlev > cv.depth && finite3(target)) ? 1u : 0u;
        }
    }
    float4 col = make_float4(value.x, value.y, value.z, 1.0f);
    if (isnan3(value)) col = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    A.intermediate[pix] = col;
    A.auxiliary[pix] = make_float4(st.x, st.y, st.z, 1.0f);
    const bool acc = li.accumulate != 0;
    const float count = (float)li.samples;
    if (nlev > 0 && !(nlev == 1 && (sw & kEnvMiss))) {
        // the reference traces this visibility ray a second time: W *= 1 - occluded
        float4 &w = A.cur.part(2, pix);
        w.w = B.occ[pix] ? 0.0f : w.w * (1.0f - 0.0f);
    } else {  // no reservoir written: primary miss or no surface
        A.cur.reset(pix);
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        li.normal[pix] = accumulate4(li.normal[pix], zero, acc, count);
        li.albedo[pix] = accumulate4(li.albedo[pix], zero, acc, count);
        li.position[pix] = zero;
    }
}

// Temporal pass (__raygen__temporal, amadeus_restir.cu:213-291), split around
// its two visibility rays as the spatial pass is: temporal_rays_kernel casts
// the current sample's ray (slot 0) and the reprojected previous sample's
// (slot 1) from the pixel's surface into the list, a walk answers them, and
// temporal_resolve_kernel merges with those answers.  The previous
// reservoir's M is capped in place in the first kernel, as the reference
// does before its ray.
__global__ __launch_bounds__(kSpatialBlock) void temporal_rays_kernel(RestirArgs A, SpatialBufs B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px = 0, py = 0, pix = 0;
    bool live = item < tiles.items && tiles.pixel(item, px, py);
    if (live) {
        pix = py * li.width + px;
        live = A.cur.M(pix) > 0;
    }
    f3 x = mk(0, 0, 0), d0 = mk(0, 0, 0), d1 = mk(0, 0, 0);
    float t0 = 0.0f, t1 = 0.0f;
    bool v1 = false;
    bool v0 = false;
    if (live) {
        const SurfaceHit hit = gbuffer_hit(A, pix);
        x = hit.x;
        const Resv cr = A.cur.load(pix);
        f3 D = cr.data.point - x;
        float d = length(D);
        d0 = D / d;
        t0 = d - kShadowEps;
        v0 = !moot_ray(cr.W, direct_li(hit, cr.data.Le, cr.data.normal, cr.data.type, d0, d));
        const int pindex = reproject(li, (int)pix, x);
        if (pindex >= 0) {
            const int pm = A.prev.M((uint32_t)pindex);
            if (pm > 0) {
                A.prev.store_M((uint32_t)pindex, min(pm, A.m_cap));
                const float4 p = A.prev.part(2, (uint32_t)pindex);
                D = mk(p.x, p.y, p.z) - x;
                d = length(D);
                d1 = D / d;
                t1 = d - kShadowEps;
                v1 = true;
            }
        }
    }
    emit_ray(B.list, &nray, v0, x, d0, t0, pix);
    emit_ray(B.list, &nray, v1, x, d1, t1, npix + pix);
    publish_rays(B.list, &nray);
}

__global__ __launch_bounds__(kSpatialBlock) void temporal_resolve_kernel(RestirArgs A, SpatialBufs B) {
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const Resv cr = A.cur.load(pix);
    if (cr.M <= 0) return;
    const SurfaceHit hit = gbuffer_hit(A, pix);
    const int pindex = reproject(li, (int)pix, hit.x);
    Resv merged;
    merged.data.Le = merged.data.normal = merged.data.point = mk(0, 0, 0);
    merged.data.type = 0;
    merged.w = merged.W = 0.0f;
    merged.M = 0;
    merged.seed = ld3(A.auxiliary[pix]);
    // occluded_target: 0 when the ray was blocked, else the unshadowed sum
    f3 D = cr.data.point - hit.x;
    float d = length(D);
    const f3 Lu = direct_li(hit, cr.data.Le, cr.data.normal, cr.data.type, D / d, d);
    f3 Li = (!moot_ray(cr.W, Lu) && B.occ[pix]) ? mk(0.0f, 0.0f, 0.0f) : Lu;
    const float t_current = Li.x + Li.y + Li.z;
    merged.update(cr.data, t_current * cr.W * (float)cr.M);
    int N = cr.M;
    if (pindex >= 0) {
        Resv pr = A.prev.load((uint32_t)pindex);
        if (pr.M > 0) {
            pr.M = min(pr.M, A.m_cap);
            D = pr.data.point - hit.x;
            d = length(D);
            Li = B.occ[npix + pix] ? mk(0.0f, 0.0f, 0.0f)
                                   : direct_li(hit, pr.data.Le, pr.data.normal, pr.data.type, D / d, d);
            const float t_prev = Li.x + Li.y + Li.z;
            merged.update(pr.data, t_prev * pr.W * (float)pr.M);
            N += (t_prev > 0.0f ? 1 : 0) * pr.M;
        }
    }
    merged.M = N;
    merged.resample(target(merged.data, hit));
    A.cur.store(pix, merged);
}


// Neighbour picks -> the resampling-target rays (origin: this pixel's
// surface, towards the neighbour's sample).
__global__ __launch_bounds__(kSpatialBlock) void spatial_rays_kernel(RestirArgs A, SpatialBufs B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px = 0, py = 0, pix = 0;
    bool live = item < tiles.items && tiles.pixel(item, px, py);
    if (live) {
        pix = py * li.width + px;
        live = A.cur.M(pix) > 0;
    }
    f3 x = mk(0, 0, 0), rs = mk(0, 0, 0);
    SurfaceHit hit;
    if (live) {
        hit = gbuffer_hit(A, pix);
        x = hit.x;
        rs = ld3(A.auxiliary[pix]);
        rand_uniform(rs);  // the current reservoir's update
    }
    for (int i = 0; i < A.neighbours; i++) {
        bool valid = false;
        f3 dir = mk(0, 0, 0);
        float tmax = 0.0f;
        if (live) {
            const uint32_t i0 = spatial_neighbour_px(li, pix, rs, A.radius);
            const Resv nb = A.cur.load(i0);
            if (nb.M > 0) {
                rand_uniform(rs);  // that neighbour's update
                const f3 Dn = nb.data.point - x;
                const float dn = length(Dn);
                dir = Dn / dn;
                tmax = dn - kShadowEps;
                valid = !moot_ray(nb.W, direct_li(hit, nb.data.Le, nb.data.normal, nb.data.type, dir, dn));
            }
        }
        emit_ray(B.list, &nray, valid, x, dir, tmax, (uint32_t)i * npix + pix);
    }
    publish_rays(B.list, &nray);
}

// Resampling over the neighbours with the traced visibilities; emits the 1/Z
// rays (from each non-empty neighbour's surface towards the merged sample)
// and, when the merged sample's visibility from this pixel is not already
// known, the shading ray.  The merged reservoir (W still open) goes to the
// destination planes.
__global__ __launch_bounds__(kSpatialBlock) void spatial_merge_kernel(RestirArgs A, SpatialBufs B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    const int S = A.neighbours;
    uint32_t px = 0, py = 0, pix = 0;
    bool live = item < tiles.items && tiles.pixel(item, px, py);
    if (live) {
        pix = py * li.width + px;
        live = A.cur.M(pix) > 0;
    }
    Resv merged;
    merged.data.Le = merged.data.normal = merged.data.point = mk(0, 0, 0);
    merged.data.type = 0;
    merged.w = merged.W = 0.0f;
    merged.M = 0;
    merged.seed = mk(0, 0, 0);
    SurfaceHit hit;
    bool sel_known = false, sel_occ = false, z_moot = false, final_moot = false;
    float t_merged = 0.0f;
    if (live) {
        hit = gbuffer_hit(A, pix);
        const Resv current = A.cur.load(pix);
        merged.seed = ld3(A.auxiliary[pix]);
        const float t_current = target(current.data, hit);
        merged.update(current.data, t_current * current.W * (float)current.M);
        int cnt = 0;
        for (int i = 0; i < S; i++) {
            const uint32_t i0 = spatial_neighbour_px(li, pix, merged.seed, A.radius);
            const Resv nb = A.cur.load(i0);
            if (nb.M > 0) {
                const f3 Dn = nb.data.point - hit.x;
                const float dn = length(Dn);
                const f3 Lu = direct_li(hit, nb.data.Le, nb.data.normal, nb.data.type, Dn / dn, dn);
                const bool occ = !moot_ray(nb.W, Lu) && B.occ[(size_t)i * npix + pix] != 0;
                const f3 Ln = occ ? mk(0.0f, 0.0f, 0.0f) : Lu;
                const float t_nb = Ln.x + Ln.y + Ln.z;
                if (merged.update(nb.data, t_nb * nb.W * (float)nb.M)) {
                    sel_known = true;
                    sel_occ = occ;
                }
            }
            cnt += nb.M;
        }
        merged.M = current.M + cnt;
        A.prev.store(pix, merged);
        // W = w / (target * Z) is 0 whatever Z when w = 0 or the target is
        // not positive: then no 1/Z ray can change it, nor the shading ray
        // (finite lighting x 0)
        t_merged = target(merged.data, hit);
        z_moot = merged.w == 0.0f || !(t_merged > 0.0f);
        if (z_moot && !sel_known) {
            const f3 D = merged.data.point - hit.x;
            const float d = length(D);
            const f3 Lf = direct_li(hit, merged.data.Le, merged.data.normal, merged.data.type, D / d, d);
            final_moot = isfinite(Lf.x) && isfinite(Lf.y) && isfinite(Lf.z);
        }
        B.flags[pix] = (uint8_t)((sel_known ? 1 : 0) | (sel_occ ? 2 : 0) | (final_moot ? 4 : 0));
    }
    // 1/Z rays: replay the picks from the auxiliary seed
    f3 rs = mk(0, 0, 0);
    if (live) {
        rs = ld3(A.auxiliary[pix]);
        rand_uniform(rs);
    }
    for (int i = 0; i < S; i++) {
        bool valid = false;
        f3 o = mk(0, 0, 0), dir = mk(0, 0, 0);
        float tmax = 0.0f;
        if (live) {
            const uint32_t i0 = spatial_neighbour_px(li, pix, rs, A.radius);
            const int m = A.cur.M(i0);
            int zm = 0;
            if (m > 0) {  // `entering` carries neighbour validity
                rand_uniform(rs);
                SurfaceHit nh;
                nh.mat = gbuffer_material(A, i0);
                nh.entering = true;
                const float4 nn = li.normal[i0], xx = li.position[i0];
                nh.n = mk(nn.x, nn.y, nn.z);
                nh.x = mk(xx.x, xx.y, xx.z);
                nh.wo = hit.wo;
                const f3 D = merged.data.point - nh.x;
                const float d = length(D);
                const f3 Li = direct_li(nh, merged.data.Le, merged.data.normal, merged.data.type, D / d, d);
                zm = (!z_moot && Li.x + Li.y + Li.z > 0.0f) ? m : 0;
                o = nh.x;
                dir = D / d;
                tmax = d - kShadowEps;
                valid = zm > 0;  // the ray counts only towards a positive target
            }
            B.zm[(size_t)i * npix + pix] = zm;
        }
        emit_ray(B.list, &nray, valid, o, dir, tmax, (uint32_t)i * npix + pix);
    }
    bool valid = false;
    f3 dir = mk(0, 0, 0);
    float tmax = 0.0f;
    if (live && !sel_known && !final_moot) {
        const f3 D = merged.data.point - hit.x;
        const float d = length(D);
        dir = D / d;
        tmax = d - kShadowEps;
        valid = true;
    }
    emit_ray(B.list, &nray, valid, hit.x, dir, tmax, (uint32_t)S * npix + pix);
    publish_rays(B.list, &nray);
}

// Z from the 1/Z visibilities, W, shading with the merged sample, colour.
__global__ __launch_bounds__(kSpatialBlock) void spatial_resolve_kernel(RestirArgs A, SpatialBufs B) {
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const int S = A.neighbours;
    const int Mc = A.cur.M(pix);
    f3 direct = mk(0, 0, 0);
    if (Mc > 0) {
        const SurfaceHit hit = gbuffer_hit(A, pix);
        Resv merged = A.prev.load(pix);
        int Z = Mc;
        for (int i = 0; i < S; i++) {
            const size_t k = (size_t)i * npix + pix;
            const int zm = B.zm[k];
            if (zm > 0 && B.occ[k] == 0) Z += zm;
        }
        const float t_merged = target(merged.data, hit);
        const float denominator = t_merged * (float)Z;
        merged.W = (denominator > 0.0f) ? merged.w / denominator : 0.0f;
        const f3 D = merged.data.point - hit.x;
        const float d = length(D);
        const uint8_t fl = B.flags[pix];
        const bool occ = (fl & 1) ? (fl & 2) != 0 : ((fl & 4) ? false : B.occ[(size_t)S * npix + pix] != 0);
        const f3 lighting = occ ? mk(0.0f, 0.0f, 0.0f)
                                : direct_li(hit, merged.data.Le, merged.data.normal, merged.data.type, D / d, d);
        direct = hit.mat.emission + lighting * merged.W;
        A.prev.store(pix, merged);
    } else {
        A.prev.reset(pix);
    }
    const float4 im = A.intermediate[pix];
    float4 col = make_float4(direct.x + im.x, direct.y + im.y, direct.z + im.z, 1.0f + im.w);
    if (isnan(col.x) || isnan(col.y) || isnan(col.z) || isnan(col.w)) col = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    li.color[pix] = accumulate4(li.color[pix], col, li.accumulate != 0, (float)li.samples);
}

__global__ void seed_kernel(Planes a, Planes b, uint32_t w, uint32_t h, uint32_t user) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    Resv r;
    r.data.Le = r.data.normal = r.data.point = mk(0, 0, 0);
    r.data.type = 0;
    r.w = r.W = 0.0f;
    r.M = 0;
    r.seed = reservoir_seed(i % w, i / w, user);
    a.store(i, r);
    b.store(i, r);
}

// Reservoir planes -> 16 floats per pixel {Le, normal, point, type, w, W, M, seed}
__global__ void unpack_kernel(Planes a, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Resv r = a.load(i);
    float *o = out + 16 * (size_t)i;
    o[0] = r.data.Le.x; o[1] = r.data.Le.y; o[2] = r.data.Le.z;
    o[3] = r.data.normal.x; o[4] = r.data.normal.y; o[5] = r.data.normal.z;
    o[6] = r.data.point.x; o[7] = r.data.point.y; o[8] = r.data.point.z;
    o[9] = (float)r.data.type;
    o[10] = r.w; o[11] = r.W; o[12] = (float)r.M;
    o[13] = r.seed.x; o[14] = r.seed.y; o[15] = r.seed.z;
}

}  // namespace

void restir_read(Context &ctx, RestirState &st, int which, float *host_out) {
    KRT_REQUIRE(st.w > 0, KRT_ERR_STATE, "ReSTIR buffers not loaded");
    uint32_t n = st.w * st.h;
    int b = which == 0 ? 1 - st.cur : st.cur;  // 0: the reservoirs the next frame merges with
    DevBuf tmp(sizeof(float) * 16 * (size_t)n);
    unpack_kernel<<<(n + 255) / 256, 256, 0, ctx.stream>>>(Planes{st.res[b].as<float4>(), n}, n, tmp.as<float>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipMemcpyAsync(host_out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost, ctx.stream));
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

void restir_alloc(Context &ctx, RestirState &st, uint32_t w, uint32_t h, uint32_t user_seed) {
    size_t n = (size_t)w * h;
    st.w = w;
    st.h = h;
    st.res[0].alloc(sizeof(float4) * 4 * n);
    st.res[1].alloc(sizeof(float4) * 4 * n);
    st.cur = 0;
    st.gmat.alloc(sizeof(uint32_t) * n);
    st.guv.alloc(sizeof(float2) * n);
    st.intermediate.alloc(sizeof(float4) * n);
    st.auxiliary.alloc(sizeof(float4) * n);
    KRT_HIP(hipMemsetAsync(st.gmat.p, 0, st.gmat.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(st.guv.p, 0, st.guv.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(st.intermediate.p, 0, st.intermediate.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(st.auxiliary.p, 0, st.auxiliary.bytes, ctx.stream));
    Planes a{st.res[0].as<float4>(), (uint32_t)n}, b{st.res[1].as<float4>(), (uint32_t)n};
    seed_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx.stream>>>(a, b, w, h, user_seed);
    KRT_HIP(hipGetLastError());
    for (int k = 0; k < 16; k++) st.prev_pv[k] = 0.0f;
}

void restir_render(Context &ctx, Scene &s, const LaunchInfo &li_in, const RestirOptions &opt, RestirState &st,
                   unsigned long long *ctr, PassTimes &times, const CacheOptions *copt, CacheState *cs, bool sync) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(st.w == li_in.width && st.h == li_in.height, KRT_ERR_STATE, "ReSTIR buffers not loaded");
    KRT_REQUIRE(opt.neighbours >= 0 && opt.neighbours <= kMaxNeighbours, KRT_ERR_INVALID,
                "restir neighbours must be in [0, 8]");
    KRT_REQUIRE(opt.candidates >= 0 && opt.spatial_passes >= 1 && opt.m_cap >= 0 && opt.radius >= 0,
                KRT_ERR_INVALID, "restir options must be non-negative and spatial_passes >= 1");
    LaunchInfo li = li_in;
    for (int k = 0; k < 16; k++) li.prev_pv[k] = st.prev_pv[k];
    const uint32_t npix = li.width * li.height;
    RestirArgs A;
    A.li = li;
    A.sv = make_view(s);
    A.gmat = st.gmat.as<uint32_t>();
    A.guv = st.guv.as<float2>();
    A.intermediate = st.intermediate.as<float4>();
    A.auxiliary = st.auxiliary.as<float4>();
    // producer / consumer workgroups (= list segments) and the walks' spill
    // slab (16-entry LDS stacks, up to two roles of kMaxResidentBlocks)
    const uint32_t items = ((li.width + 7u) / 8u) * ((li.height + 7u) / 8u) * 64u;
    const uint32_t gb = (items + kSpatialBlock - 1) / kSpatialBlock;
    A.spill = spill_buffer(ctx, s, 2 * kMaxResidentBlocks * kTraceBlock, kLdsDepthLean);
    A.ctr = ctr;
    A.candidates = opt.candidates;
    A.m_cap = opt.m_cap;
    A.neighbours = opt.neighbours;
    A.radius = opt.radius;
    const bool cached = copt && cs && copt->enabled;
    KRT_REQUIRE(sync || !cached, KRT_ERR_INVALID, "the radiance cache needs a synchronous ReSTIR frame");
    if (cached) {
        KRT_REQUIRE(copt->depth >= 1, KRT_ERR_INVALID, "restir radiance cache needs cache_depth >= 1");
        A.cv = cache_begin(ctx, s, *copt, *cs, npix, li.user_seed);
    }
    auto planes = [&](int b) { return Planes{st.res[b].as<float4>(), npix}; };
    hipEvent_t ev[16];
    int nev = 0;
    auto mark = [&]() {
        if (!sync) return;  // untimed: queued on ctx.stream, no host wait
        KRT_HIP(hipEventCreate(&ev[nev]));
        KRT_HIP(hipEventRecord(ev[nev], ctx.stream));
        nev++;
    };
    // shadow-ray lists of the passes: segments sized for the largest per-lane
    // count (1 in the initial pass, 2 temporal rays, S + 1 in the spatial
    // merge); the initial pass's closest rays have a list of their own
    const int S = opt.neighbours;
    const uint32_t per_lane = (uint32_t)std::max(S + 1, 2);
    const size_t list_slots = (size_t)gb * kSpatialBlock * per_lane;
    const size_t slots = (size_t)per_lane * npix;
    if (st.rays.bytes < sizeof(float4) * 2 * list_slots) st.rays.alloc(sizeof(float4) * 2 * list_slots);
    if (st.seg_counts.bytes < sizeof(uint32_t) * gb) st.seg_counts.alloc(sizeof(uint32_t) * gb);
    if (st.seg_prefix.bytes < sizeof(uint32_t) * (gb + 1)) st.seg_prefix.alloc(sizeof(uint32_t) * (gb + 1));
    if (st.dense.bytes < sizeof(float4) * 2 * list_slots) st.dense.alloc(sizeof(float4) * 2 * list_slots);
    if (st.occ.bytes < slots) st.occ.alloc(slots);
    if (st.zm.bytes < sizeof(int32_t) * ((size_t)S * npix + 1)) st.zm.alloc(sizeof(int32_t) * ((size_t)S * npix + 1));
    if (st.flags.bytes < npix) st.flags.alloc(npix);
    SpatialBufs B;
    B.list.rays = st.rays.as<float4>();
    B.list.counts = st.seg_counts.as<uint32_t>();
    B.list.prefix = st.seg_prefix.as<uint32_t>();
    B.list.dense = st.dense.as<float4>();
    B.list.nseg = gb;
    B.occ = st.occ.as<uint8_t>();
    B.zm = st.zm.as<int32_t>();
    B.flags = st.flags.as<uint8_t>();
    int *lean_spill = A.spill;  // sized above for the list walks too
    // With the cache, the inference reads only the query records (all written
    // by the cache-vertex shading pass) and runs on the side stream while the
    // main stream walks the training paths' remaining levels and the temporal
    // pass; the fold into the intermediate buffer follows the initial pass's
    // compose there, and the first reader of that buffer waits for it.
    hipStream_t side = ctx.side ? ctx.side : ctx.stream;
    hipEvent_t queried = nullptr, inf0 = nullptr, inf1 = nullptr, folded = nullptr;
    if (cached) {
        KRT_HIP(hipEventCreate(&queried));
        KRT_HIP(hipEventCreate(&inf0));
        KRT_HIP(hipEventCreate(&inf1));
        KRT_HIP(hipEventCreate(&folded));
    }
    bool inferring = false;
    auto start_infer = [&]() {
        KRT_HIP(hipEventRecord(queried, ctx.stream));
        KRT_HIP(hipStreamWaitEvent(side, queried, 0));
        KRT_HIP(hipEventRecord(inf0, side));
        cache_infer(*cs, npix, side);
        KRT_HIP(hipEventRecord(inf1, side));
        inferring = true;
    };
    mark();
    A.cur = planes(st.cur);
    A.prev = planes(1 - st.cur);
    {
        const size_t levels = (size_t)li.max_depth + 1;
        const size_t cl_slots = (size_t)gb * kSpatialBlock;
        if (st.wf_hits.bytes < sizeof(float4) * npix) st.wf_hits.alloc(sizeof(float4) * npix);
        if (s.instanced && st.wf_inst.bytes < sizeof(int) * npix) st.wf_inst.alloc(sizeof(int) * npix);
        if (st.wf_state.bytes < sizeof(float4) * npix) st.wf_state.alloc(sizeof(float4) * npix);
        if (st.wf_hist.bytes < sizeof(float4) * 3 * levels * npix) st.wf_hist.alloc(sizeof(float4) * 3 * levels * npix);
        if (st.wf_occ.bytes < levels * npix) st.wf_occ.alloc(levels * npix);
        if (cached && st.wf_lec.bytes < sizeof(float4) * npix) st.wf_lec.alloc(sizeof(float4) * npix);
        if (st.cl_rays.bytes < sizeof(float4) * 2 * cl_slots) st.cl_rays.alloc(sizeof(float4) * 2 * cl_slots);
        if (st.cl_counts.bytes < sizeof(uint32_t) * gb) st.cl_counts.alloc(sizeof(uint32_t) * gb);
        if (st.cl_prefix.bytes < sizeof(uint32_t) * (gb + 1)) st.cl_prefix.alloc(sizeof(uint32_t) * (gb + 1));
        if (st.cl_dense.bytes < sizeof(float4) * 2 * cl_slots) st.cl_dense.alloc(sizeof(float4) * 2 * cl_slots);
        InitWave W;
        W.cl = RayList{st.cl_rays.as<float4>(), st.cl_counts.as<uint32_t>(), (uint32_t)kSpatialBlock, gb,
                       st.cl_prefix.as<uint32_t>(), st.cl_dense.as<float4>()};
        W.sl = B.list;
        W.sl.seg = kSpatialBlock;
        W.hits = st.wf_hits.as<float4>();
        W.hit_inst = s.instanced ? st.wf_inst.as<int>() : nullptr;
        W.state = st.wf_state.as<float4>();
        W.hist = st.wf_hist.as<float4>();
        W.occ = st.wf_occ.as<uint8_t>();
        W.lec = st.wf_lec.as<float4>();
        init_camera_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, W);
        KRT_HIP(hipGetLastError());
        walk_closest_list(ctx, s.instanced, A.sv, W.cl, W.hits, W.hit_inst, lean_spill, ctr, 0);
        // Past the cache vertex only the training paths (at most cv.M) go on, so
        // those levels' shading grids, and the segments of the lists they emit,
        // shrink to the training paths' count; with no training paths they end.
        const uint32_t train_blocks = (A.cv.M + kSpatialBlock - 1) / kSpatialBlock;
        uint32_t in_seg = gb;  // segments of the closest list the next shading pass reads
        for (int level = 0; level <= li.max_depth; level++) {
            uint32_t grid = gb;
            if (cached && level > A.cv.depth) {
                if (train_blocks == 0) break;
                grid = std::min(gb, train_blocks);
            }
            InitWave Wl = W;
            Wl.cl.nseg = in_seg;
            if (s.instanced) init_shade_kernel<true><<<grid, kSpatialBlock, 0, ctx.stream>>>(A, Wl, level);
            else init_shade_kernel<false><<<grid, kSpatialBlock, 0, ctx.stream>>>(A, Wl, level);
            KRT_HIP(hipGetLastError());
            if (cached && level == A.cv.depth) start_infer();
            Wl.cl.nseg = Wl.sl.nseg = in_seg = grid;
            // the level's shadow rays and the next level's closest rays
            if (level < li.max_depth)
                walk_pair(ctx, s.instanced, A.sv, Wl.cl, Wl.hits, Wl.hit_inst, Wl.sl, Wl.occ + (size_t)level * npix,
                          lean_spill, ctr, 0);
            else
                walk_shadow_list(ctx, s.instanced, A.sv, Wl.sl, Wl.occ + (size_t)level * npix, lean_spill, ctr, 0);
        }
        init_compose_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, W);
        KRT_HIP(hipGetLastError());
    }
    mark();
    if (cached) {
        if (!inferring) start_infer();  // the cache vertex lies beyond the path's depth
        KRT_HIP(hipStreamWaitEvent(side, ev[nev - 1], 0));
        cache_fold_intermediate(*cs, npix, st.intermediate.as<float4>(), side);
        KRT_HIP(hipEventRecord(folded, side));
    }
    B.list.seg = kSpatialBlock * 2u;
    temporal_rays_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, B);
    KRT_HIP(hipGetLastError());
    walk_shadow_list(ctx, s.instanced, A.sv, B.list, B.occ, lean_spill, ctr, 1);
    temporal_resolve_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, B);
    KRT_HIP(hipGetLastError());
    mark();
    if (opt.spatial_passes > 0) {
        for (int pass = 0; pass < opt.spatial_passes; pass++) {
            A.cur = planes(st.cur);
            A.prev = planes(1 - st.cur);
            B.list.seg = kSpatialBlock * (uint32_t)S;
            spatial_rays_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            walk_shadow_list(ctx, s.instanced, A.sv, B.list, B.occ, lean_spill, ctr, 2);
            B.list.seg = kSpatialBlock * (uint32_t)(S + 1);
            spatial_merge_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            walk_shadow_list(ctx, s.instanced, A.sv, B.list, B.occ, lean_spill, ctr, 2);
            if (cached && pass == 0) KRT_HIP(hipStreamWaitEvent(ctx.stream, folded, 0));
            spatial_resolve_kernel<<<gb, kSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            st.cur = 1 - st.cur;  // swap(current, previous), restir.cuh:294
        }
    }
    mark();
    if (cached) {
        if (opt.spatial_passes == 0) KRT_HIP(hipStreamWaitEvent(ctx.stream, folded, 0));
        cache_train(ctx, *cs, A.cv);
    }
    mark();
    st.cur = 1 - st.cur;  // final swap, restir.cuh:300
    for (int k = 0; k < 16; k++) st.prev_pv[k] = li.pv[k];
    if (!sync) {
        for (hipEvent_t e : {queried, inf0, inf1, folded})
            if (e) (void)hipEventDestroy(e);
        return;
    }
    KRT_HIP(hipEventSynchronize(ev[nev - 1]));
    // segments: initial, temporal, spatial, training; the inference is timed
    // on the side stream
    float seg[4] = {0, 0, 0, 0};
    for (int k = 0; k + 1 < nev && k < 4; k++) KRT_HIP(hipEventElapsedTime(&seg[k], ev[k], ev[k + 1]));
    times.ms[0] = seg[0];
    times.ms[1] = seg[1];
    times.ms[2] = seg[2];
    times.ms[3] = 0.0f;
    times.ms[4] = seg[3];
    if (cached) KRT_HIP(hipEventElapsedTime(&times.ms[3], inf0, inf1));
    if (nev > 1) KRT_HIP(hipEventElapsedTime(&times.frame, ev[0], ev[nev - 1]));
    for (int k = 0; k < nev; k++) (void)hipEventDestroy(ev[k]);
    for (hipEvent_t e : {queried, inf0, inf1, folded})
        if (e) (void)hipEventDestroy(e);
}

}  // namespace krt

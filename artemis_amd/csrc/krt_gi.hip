// ReSTIR GI attachment (no reference implementation: the reference has no GI
// resampling; the algorithm follows Ouyang et al. 2021, "ReSTIR GI: Path
// Resampling for Real-Time Path Tracing").
//
//  initial : primary hit x_v (fixed -0.5 pixel offset, as the ReSTIR raygen of
//            source/optix/amadeus_restir.cu:31-65), direct = Le + Ld (the
//            path tracer's NEE), one BSDF-sampled direction wi; its hit x_s =
//            x_v + t*wi and the outgoing radiance L_o = the path tracer's value
//            of the rest of the path (levels 1..max_depth) form the sample.
//            p_hat = sum(L_o) (0 when x_s lies below x_v's tangent plane),
//            w = p_hat / pdf(wi).
//  temporal: merge with last frame's reservoir (same pixel while
//            accumulating, reprojected otherwise; normal/depth similarity),
//            M capped, reconnection Jacobian, visibility x_v -> x_s.
//  spatial : `neighbours` reservoirs in a `radius` disk, same tests, 1/Z bias
//            correction; the last pass shades
//            direct + f(x_v, wo, wi_s) * L_o * |n_v . wi_s| * W.
// Reservoirs: five float4 planes {x_s, M}, {n_s, w}, {L_o, W}, {x_owner}, {seed}.
#include <hip/hip_runtime.h>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_shadow.h"
#include "krt_traverse.h"

namespace krt {

namespace {

constexpr float kGiMaxJacobian = 10.0f;
constexpr float kGiNormalCos = 0.9f;
constexpr float kGiDepthRel = 0.1f;
constexpr int kGiMaxNeighbours = 8;

struct GSample {
    f3 xs, ns, Lo;
};

struct GRes {
    GSample s;
    f3 xo;
    float w, W;
    int M;
    f3 seed;
    KRT_DEV bool update(const GSample &x, float weight) {
        w += weight;
        float eta = rand_uniform(seed);
        bool sel = (eta * w) < weight;
        if (sel) s = x;
        M++;
        return sel;
    }
};

// A frame's GI reservoirs: per pixel one 64 B record {xs, M}, {ns, w},
// {Lo, W}, {xo} (a neighbour's sample is one cache line to fetch), then the
// seeds as a plane of their own after the n records.
struct GPlanes {
    float4 *p;
    uint32_t n;
    KRT_DEV float4 &part(uint32_t k, uint32_t i) const { return p[4 * (size_t)i + k]; }
    KRT_DEV float4 &seed(uint32_t i) const { return p[4 * (size_t)n + i]; }
    KRT_DEV GRes load(uint32_t i) const {
        float4 a = part(0, i), b = part(1, i), c = part(2, i), d = part(3, i), e = seed(i);
        GRes r;
        r.s.xs = mk(a.x, a.y, a.z);
        r.M = __float_as_int(a.w);
        r.s.ns = mk(b.x, b.y, b.z);
        r.w = b.w;
        r.s.Lo = mk(c.x, c.y, c.z);
        r.W = c.w;
        r.xo = mk(d.x, d.y, d.z);
        r.seed = mk(e.x, e.y, e.z);
        return r;
    }
    KRT_DEV void store(uint32_t i, const GRes &r) const {
        part(0, i) = make_float4(r.s.xs.x, r.s.xs.y, r.s.xs.z, __int_as_float(r.M));
        part(1, i) = make_float4(r.s.ns.x, r.s.ns.y, r.s.ns.z, r.w);
        part(2, i) = make_float4(r.s.Lo.x, r.s.Lo.y, r.s.Lo.z, r.W);
        part(3, i) = make_float4(r.xo.x, r.xo.y, r.xo.z, 0.0f);
        seed(i) = make_float4(r.seed.x, r.seed.y, r.seed.z, 0.0f);
    }
    KRT_DEV void reset(uint32_t i) const {
        part(0, i).w = __int_as_float(0);
        part(1, i).w = 0.0f;
        part(2, i).w = 0.0f;
    }
    KRT_DEV void store_M(uint32_t i, int M) const { part(0, i).w = __int_as_float(M); }
    KRT_DEV int M(uint32_t i) const { return __float_as_int(part(0, i).w); }
};

// G-buffer: {x, depth}, {n, valid}, material id, texture coordinates
struct GBuf {
    float4 *xd, *nv;
    uint32_t *mat;
    float2 *uv;
};

struct GPix {
    bool valid;
    f3 x, n;
    float depth;
    uint32_t mat;
};

KRT_DEV GPix load_g(const GBuf &g, uint32_t i) {
    float4 a = g.xd[i], b = g.nv[i];
    GPix p;
    p.x = mk(a.x, a.y, a.z);
    p.depth = a.w;
    p.n = mk(b.x, b.y, b.z);
    p.valid = b.w != 0.0f;
    p.mat = g.mat[i];
    return p;
}

KRT_DEV float gi_target(const GSample &s, f3 xv, f3 nv) {
    f3 D = s.xs - xv;
    if (!(dot(nv, D) > 0.0f)) return 0.0f;
    return s.Lo.x + s.Lo.y + s.Lo.z;
}

KRT_DEV float gi_jacobian(f3 xq, f3 xr, const GSample &s) {
    f3 dq = xq - s.xs, dr = xr - s.xs;
    float lq2 = dot(dq, dq), lr2 = dot(dr, dr);
    if (!(lq2 > 0.0f) || !(lr2 > 0.0f)) return 0.0f;
    float cq = fabsf(dot(s.ns, dq)) / sqrtf(lq2);
    float cr = fabsf(dot(s.ns, dr)) / sqrtf(lr2);
    if (!(cq > 0.0f)) return 0.0f;
    float J = (cr / cq) * (lq2 / lr2);
    return (J > kGiMaxJacobian) ? 0.0f : J;
}

KRT_DEV bool gi_similar(const GPix &a, const GPix &b) {
    return b.valid && dot(a.n, b.n) > kGiNormalCos && fabsf(a.depth - b.depth) < kGiDepthRel * a.depth;
}

struct GiArgs {
    LaunchInfo li;
    SceneView sv;
    GPlanes cur, prev;
    GBuf g, pg;
    float4 *direct, *aux;
    int *spill;
    unsigned long long *ctr;
    int m_cap, neighbours, radius, shade;
};

KRT_DEV bool gi_visible(const SceneView &sv, f3 xv, f3 xs, const TraceStack &stk, TravStats &ts, RayCount &rc) {
    f3 D = xs - xv;
    float d = length(D);
    rc.shadow++;
    return !occluded(sv, xv, D / d, d - kShadowEps, stk, ts);
}

KRT_DEV void gi_shade(const GiArgs &A, uint32_t pix, const GPix &gb, const GRes &R) {
    const LaunchInfo &li = A.li;
    float4 dv = A.direct[pix];
    f3 col = mk(dv.x, dv.y, dv.z);
    if (gb.valid && R.W > 0.0f) {
        f3 wi = normalize(R.s.xs - gb.x);
        SurfaceHit sh;
        sh.mat = load_material(A.sv.mats, gb.mat);
        const uint4 tx = material_tex(A.sv.mats, gb.mat);
        if (tx.w) {  // the primary surface's textured material (the oracle's G-buffer holds it resolved)
            const float2 uv = A.g.uv[pix];
            apply_textures(A.sv.tex, tx, uv.x, uv.y, sh.mat);
        }
        sh.entering = true;
        sh.n = gb.n;
        sh.wo = normalize(ld3(li.center) - gb.x);
        sh.x = gb.x;
        f3 f = brdf(sh, wi, eDiffuse);
        col = col + (f * R.s.Lo * fabsf(dot(gb.n, wi))) * R.W;
    }
    float4 c4 = make_float4(col.x, col.y, col.z, 1.0f);
    if (isnan3(col)) c4 = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
    li.color[pix] = accumulate4(li.color[pix], c4, li.accumulate != 0, (float)li.samples);
}

// The initial pass in wavefront form, split at every ray as the path
// tracer is.  Level 0 is the primary hit x_v (direct
// = Le + Ld, the BSDF sample wi_0); levels 1..max_depth are the secondary
// path, whose value is the sample's L_o.  Per level: a closest walk of the
// level's rays, gi_shade_kernel (Ld's light sample and shadow ray, the BSDF
// sample -> the next level's ray, the level's history), a shadow walk; then
// gi_compose_kernel folds the history and forms the sample, the reservoir and
// the G-buffer.  The same expressions in the same order (the random sequence
// included) in every launch; the oracle holds it bit for bit.
struct GiWave {
    RayList cl, sl;
    float4 *hits;   // per pixel: {t, b1, b2, prim bits}
    float4 *state;  // per pixel: {seed, secondary levels with history (int bits)}
    float4 *hist;   // per level, per pixel: {Le, has_pdf}, {T, Ld flag}, {Ld contribution unblocked, Ld pdf}
    float4 *v0;     // per pixel: {x_v, pdf0}, {n_v, material id bits}, {wi_0, flags: 1 hit0, 2 valid1}, {n_1, t_1}
    uint8_t *occ;   // per level, per pixel
};

constexpr int kGiWaveBlock = 256;

__global__ __launch_bounds__(kGiWaveBlock) void gi_camera_kernel(LaunchInfo li, GiWave B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px = 0, py = 0, pix = 0;
    const bool live = item < tiles.items && tiles.pixel(item, px, py);
    f3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    if (live) {
        pix = py * li.width + px;
        f3 seed = pixel_seed(px, py, (uint32_t)li.samples, li.user_seed);
        make_ray(li, px, py, seed, false, o, d);
        B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float(0));
        B.v0[2 * (size_t)npix + pix] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
    }
    emit_ray(B.cl, &nray, live, o, d, kTraceTmax, pix);
    publish_rays(B.cl, &nray);
}

__global__ __launch_bounds__(kGiWaveBlock) void gi_shade_kernel(LaunchInfo li, SceneView sv, GiWave B, int level) {
    __shared__ uint32_t ncl, nsl;
    if (threadIdx.x == 0) ncl = nsl = 0;
    __syncthreads();
    const uint32_t npix = li.width * li.height;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned total = sv.nl;
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
        h.inst = -1;
        Surface sf;
        if (h.prim >= 0 && load_surface(sv, h, d, sf)) {
            const Material mat = surface_material(sv, sf);
            const f3 n = sf.n;
            const f3 x = offset_point(sf.x, n, mat.type);
            const SurfaceHit sh{mat, sf.entering, n, -d, x};
            // Ld (krt_shade.h) with its shadow ray deferred
            f3 contr = mk(0, 0, 0);
            float lpdf = 0.0f;
            bool ld = false;
            if (total == 0) {
                rand_uniform(seed);
            } else {
                unsigned i = pick_light(seed, total);
                float pdf = 1.0f / (float)total;
                TriLight L = load_light(sv.lights, i);
                float light_pdf = 1.0f / L.area;
                f3 lpos = sample_area_light(L, seed);
                f3 wi = normalize(lpos - sh.x);
                float R = length(lpos - sh.x);
                f3 f = brdf(sh, wi, eDiffuse);
                float ldot = fabsf(dot(L.normal, wi));
                float geometric = ldot * fabsf(dot(sh.n, wi)) / (R * R);
                contr = f * L.intensity * geometric;
                pdf *= light_pdf;
                lpdf = pdf;
                ld = true;
                so = sh.x;
                sd = wi;
                stmax = R - kShadowEps;
                shadow = true;
            }
            int out;
            f3 wi;
            float pdf;
            f3 f = eval_bsdf(sh, wi, pdf, out, seed);
            float4 *hl = B.hist + 3 * (size_t)level * npix;
            hl[2 * (size_t)npix + pix] = make_float4(contr.x, contr.y, contr.z, lpdf);
            if (level == 0) {
                hl[pix] = make_float4(mat.emission.x, mat.emission.y, mat.emission.z, 0.0f);
                hl[npix + pix] = make_float4(0.0f, 0.0f, 0.0f, ld ? 1.0f : 0.0f);
                B.v0[pix] = make_float4(x.x, x.y, x.z, pdf);
                B.v0[npix + pix] = make_float4(n.x, n.y, n.z, __uint_as_float(sf.mat_id));
                if (sv.uv) B.v0[4 * (size_t)npix + pix] = make_float4(sf.u, sf.v, 0.0f, 0.0f);
                B.v0[2 * (size_t)npix + pix] = make_float4(wi.x, wi.y, wi.z, __int_as_float(1));
                B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float(0));
                if (pdf > 0.0f && li.max_depth >= 1) {
                    no = x;
                    nd = wi;
                    next = true;
                }
            } else {
                f3 T = f * fabsf(dot(wi, n)) / pdf;
                bool has_pdf = pdf > 0.0f;
                hl[pix] = make_float4(mat.emission.x, mat.emission.y, mat.emission.z, has_pdf ? 1.0f : 0.0f);
                hl[npix + pix] = make_float4(T.x, T.y, T.z, ld ? 1.0f : 0.0f);
                if (level == 1) {
                    float4 &w0 = B.v0[2 * (size_t)npix + pix];
                    w0.w = __int_as_float(__float_as_int(w0.w) | 2);
                    B.v0[3 * (size_t)npix + pix] = make_float4(n.x, n.y, n.z, h.t);
                }
                B.state[pix] = make_float4(seed.x, seed.y, seed.z, __int_as_float(level));
                if (has_pdf && level < li.max_depth) {
                    no = x;
                    nd = wi;
                    next = true;
                }
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

__global__ __launch_bounds__(kGiWaveBlock) void gi_compose_kernel(GiArgs A, GiWave B) {
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const float4 st = B.state[pix];
    const int nlev = __float_as_int(st.w);
    const float4 w0 = B.v0[2 * (size_t)npix + pix];
    const int flags = __float_as_int(w0.w);
    const bool hit0 = (flags & 1) != 0, valid1 = (flags & 2) != 0;
    // Le + Ld of the given level, the shadow ray's answer applied
    auto direct_of = [&](int lv) {
        const float4 *hl = B.hist + 3 * (size_t)lv * npix;
        const float4 e = hl[pix], t = hl[npix + pix], c = hl[2 * (size_t)npix + pix];
        f3 Ldv = mk(0.0f, 0.0f, 0.0f);
        if (t.w != 0.0f) {
            const f3 contr = B.occ[(size_t)lv * npix + pix] ? mk(0.0f, 0.0f, 0.0f) : mk(c.x, c.y, c.z);
            Ldv = contr / c.w;
        }
        return mk(e.x, e.y, e.z) + Ldv;
    };
    f3 direct = mk(0, 0, 0);
    if (hit0) direct = direct_of(0);
    GRes R = A.cur.load(pix);
    R.w = 0.0f;
    R.M = 0;
    R.W = 0.0f;
    float4 gxd = make_float4(0, 0, 0, 0), gnv = make_float4(0, 0, 0, 0);
    uint32_t gm = 0;
    float2 guv = make_float2(0.0f, 0.0f);
    if (hit0) {
        const float4 a = B.v0[pix], b = B.v0[npix + pix];
        const f3 xv = mk(a.x, a.y, a.z), nv = mk(b.x, b.y, b.z), wi0 = mk(w0.x, w0.y, w0.z);
        const float pdf0 = a.w;
        GSample smp;
        smp.xs = xv + wi0;
        smp.ns = -wi0;
        smp.Lo = mk(0, 0, 0);
        if (pdf0 > 0.0f && li.max_depth >= 1) {
            f3 value = mk(0, 0, 0);
            for (int k = nlev - 1; k >= 0; k--) {
                const float4 *hl = B.hist + 3 * (size_t)(k + 1) * npix;
                const float4 e = hl[pix], t = hl[npix + pix];
                f3 v = direct_of(k + 1);
                if (e.w != 0.0f) v = v + mk(t.x, t.y, t.z) * value;
                value = v;
            }
            if (valid1) {
                const float4 q = B.v0[3 * (size_t)npix + pix];
                smp.xs = xv + wi0 * q.w;
                smp.ns = mk(q.x, q.y, q.z);
                smp.Lo = value;
            }
        }
        float ph = gi_target(smp, xv, nv);
        R.update(smp, (pdf0 > 0.0f) ? ph / pdf0 : 0.0f);
        float dd = ph * (float)R.M;
        R.W = (dd > 0.0f) ? R.w / dd : 0.0f;
        R.xo = xv;
        gxd = make_float4(xv.x, xv.y, xv.z, length(xv - ld3(li.center)));
        gnv = make_float4(nv.x, nv.y, nv.z, 1.0f);
        gm = __float_as_uint(b.w);
        if (A.sv.uv) {
            const float4 q = B.v0[4 * (size_t)npix + pix];
            guv = make_float2(q.x, q.y);
        }
    }
    A.g.xd[pix] = gxd;
    A.g.nv[pix] = gnv;
    A.g.mat[pix] = gm;
    if (A.sv.uv) A.g.uv[pix] = guv;
    A.direct[pix] = make_float4(direct.x, direct.y, direct.z, 0.0f);
    A.aux[pix] = make_float4(st.x, st.y, st.z, 0.0f);
    A.cur.store(pix, R);
}

KRT_DEV int gi_reproject(const LaunchInfo &li, uint32_t pix, f3 x) {
    if (li.samples > 0) return (int)pix;
    const float *m = li.prev_pv;
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (m[0 + k] * x.x + m[4 + k] * x.y) + (m[8 + k] * x.z + m[12 + k] * 1.0f);
    float u = v[0] / v[3], vv = v[1] / v[3];
    if (!(u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f)) return -1;
    int ix = round_half_up((u + 1.0f) * 0.5f * (float)li.width);
    int iy = round_half_up((vv + 1.0f) * 0.5f * (float)li.height);
    ix = min(max(ix, 0), (int)li.width - 1);
    iy = min(max(iy, 0), (int)li.height - 1);
    return iy * (int)li.width + ix;
}

__global__ __launch_bounds__(kTraceBlock) void gi_temporal_kernel(GiArgs A) {
    __shared__ int lds_stack[kLdsDepth * kTraceBlock];
    const LaunchInfo &li = A.li;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepth, A.spill + gtid, (int)nthreads};
    TravStats ts;
    RayCount rc;
    const TileOrder tiles(li.width, li.height);
    for (uint32_t item = gtid; item < tiles.items; item += nthreads) {
        uint32_t px, py;
        if (!tiles.pixel(item, px, py)) continue;
        const uint32_t pix = py * li.width + px;
        GPix gb = load_g(A.g, pix);
        GRes R = A.cur.load(pix);
        if (gb.valid && R.M > 0) {
            GRes m;
            m.s.xs = m.s.ns = m.s.Lo = mk(0, 0, 0);
            m.xo = mk(0, 0, 0);
            m.w = m.W = 0.0f;
            m.M = 0;
            float4 ax = A.aux[pix];
            m.seed = mk(ax.x, ax.y, ax.z);
            m.update(R.s, gi_target(R.s, gb.x, gb.n) * R.W * (float)R.M);
            int N = R.M;
            int q = gi_reproject(li, pix, gb.x);
            if (q >= 0) {
                GPix pq = load_g(A.pg, (uint32_t)q);
                if (gi_similar(gb, pq)) {
                    GRes P = A.prev.load((uint32_t)q);
                    if (P.M > 0) {
                        P.M = min(P.M, A.m_cap);
                        A.prev.store_M((uint32_t)q, P.M);
                        float J = gi_jacobian(P.xo, gb.x, P.s);
                        float t = 0.0f;
                        if (J > 0.0f && gi_target(P.s, gb.x, gb.n) > 0.0f &&
                            gi_visible(A.sv, gb.x, P.s.xs, stk, ts, rc))
                            t = gi_target(P.s, gb.x, gb.n);
                        m.update(P.s, t * P.W * (float)P.M * J);
                        N += (t > 0.0f ? 1 : 0) * P.M;
                    }
                }
            }
            m.M = N;
            float tm = gi_target(m.s, gb.x, gb.n) * (float)m.M;
            m.W = (tm > 0.0f) ? m.w / tm : 0.0f;
            m.xo = gb.x;
            A.cur.store(pix, m);
            R = m;
        }
        if (A.shade) gi_shade(A, pix, gb, R);
    }
    wave_count(A.ctr ? A.ctr + 2 : nullptr, rc.shadow);
    wave_count(A.ctr ? A.ctr + 5 : nullptr, ts.nodes);
    wave_count(A.ctr ? A.ctr + 8 : nullptr, ts.tris);
}

// Spatial pass, split around its visibility rays as the ReSTIR DI spatial
// pass is: gi_spatial_rays_kernel appends the x_v -> x_s rays of the
// neighbours whose Jacobian and target are positive (slot i = neighbour i),
// a shadow-list walk answers them, gi_spatial_merge_kernel resamples with
// those answers and appends the 1/Z rays from each used neighbour's surface
// towards the merged sample (slot k = k-th used neighbour), a second walk
// answers them and gi_spatial_resolve_kernel forms Z, W and shades.  The seed
// sequence does not depend on any visibility (a pick draws three numbers, an
// accepted neighbour's update one), so every kernel replays the picks.
struct GiSpatialBufs {
    RayList list;
    uint8_t *occ;  // [slot * npix + pixel]
    int32_t *zm;   // [k * npix + pixel]: M of the k-th used neighbour when its 1/Z ray was cast, else 0
};

constexpr int kGiSpatialBlock = 256;

__global__ __launch_bounds__(kGiSpatialBlock) void gi_spatial_rays_kernel(GiArgs A, GiSpatialBufs B) {
    __shared__ uint32_t nray;
    if (threadIdx.x == 0) nray = 0;
    __syncthreads();
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px = 0, py = 0, pix = 0;
    bool live = item < tiles.items && tiles.pixel(item, px, py);
    GPix gb;
    gb.valid = false;
    gb.x = gb.n = mk(0, 0, 0);
    gb.depth = 0.0f;
    gb.mat = 0;
    f3 seed = mk(0, 0, 0);
    if (live) {
        pix = py * li.width + px;
        gb = load_g(A.g, pix);
        live = gb.valid && A.cur.M(pix) > 0;
    }
    if (live) {
        const float4 ax = A.aux[pix];
        seed = mk(ax.x, ax.y, ax.z);
        rand_uniform(seed);  // the pixel's own reservoir's update
    }
    for (int i = 0; i < A.neighbours; i++) {
        bool valid = false;
        f3 dir = mk(0, 0, 0);
        float tmax = 0.0f;
        if (live) {
            const uint32_t q = spatial_neighbour_px(li, pix, seed, A.radius);
            if (q != pix) {
                const GPix gq = load_g(A.g, q);
                if (gi_similar(gb, gq)) {
                    const GRes Nb = A.cur.load(q);
                    if (Nb.M > 0) {
                        const float J = gi_jacobian(Nb.xo, gb.x, Nb.s);
                        if (J > 0.0f && gi_target(Nb.s, gb.x, gb.n) > 0.0f) {
                            const f3 D = Nb.s.xs - gb.x;
                            const float d = length(D);
                            dir = D / d;
                            tmax = d - kShadowEps;
                            valid = true;
                        }
                        rand_uniform(seed);  // that neighbour's update
                    }
                }
            }
        }
        emit_ray(B.list, &nray, valid, gb.x, dir, tmax, (uint32_t)i * npix + pix);
    }
    publish_rays(B.list, &nray);
}

__global__ __launch_bounds__(kGiSpatialBlock) void gi_spatial_merge_kernel(GiArgs A, GiSpatialBufs B) {
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
    GPix gb;
    gb.valid = false;
    gb.x = gb.n = mk(0, 0, 0);
    gb.depth = 0.0f;
    gb.mat = 0;
    if (live) {
        pix = py * li.width + px;
        gb = load_g(A.g, pix);
        live = gb.valid && A.cur.M(pix) > 0;
    }
    GRes m;
    m.s.xs = m.s.ns = m.s.Lo = mk(0, 0, 0);
    m.xo = mk(0, 0, 0);
    m.w = m.W = 0.0f;
    m.M = 0;
    m.seed = mk(0, 0, 0);
    uint32_t used[kGiMaxNeighbours];
    int Ms[kGiMaxNeighbours];
    int nu = 0;
#pragma unroll
    for (int i = 0; i < kGiMaxNeighbours; i++) {
        used[i] = pix;
        Ms[i] = 0;
    }
    if (live) {
        const GRes R = A.cur.load(pix);
        const float4 ax = A.aux[pix];
        m.seed = mk(ax.x, ax.y, ax.z);
        m.update(R.s, gi_target(R.s, gb.x, gb.n) * R.W * (float)R.M);
        int cnt = 0;
        for (int i = 0; i < S; i++) {
            const uint32_t q = spatial_neighbour_px(li, pix, m.seed, A.radius);
            if (q == pix) continue;
            const GPix gq = load_g(A.g, q);
            if (!gi_similar(gb, gq)) continue;
            const GRes Nb = A.cur.load(q);
            if (Nb.M <= 0) continue;
            const float J = gi_jacobian(Nb.xo, gb.x, Nb.s);
            float t = 0.0f;
            if (J > 0.0f && gi_target(Nb.s, gb.x, gb.n) > 0.0f && B.occ[(size_t)i * npix + pix] == 0)
                t = gi_target(Nb.s, gb.x, gb.n);
            m.update(Nb.s, t * Nb.W * (float)Nb.M * J);
#pragma unroll
            for (int k = 0; k < kGiMaxNeighbours; k++)
                if (k == nu) {
                    used[k] = q;
                    Ms[k] = Nb.M;
                }
            nu++;
            cnt += Nb.M;
        }
        m.M = R.M + cnt;
        A.prev.store(pix, m);
    }
#pragma unroll
    for (int k = 0; k < kGiMaxNeighbours; k++) {
        if (k >= S) break;
        bool valid = false;
        f3 o = mk(0, 0, 0), dir = mk(0, 0, 0);
        float tmax = 0.0f;
        if (live) {
            int zm = 0;
            if (k < nu) {
                const GPix gq = load_g(A.g, used[k]);
                if (gi_jacobian(gq.x, gb.x, m.s) > 0.0f && gi_target(m.s, gq.x, gq.n) > 0.0f) {
                    const f3 D = m.s.xs - gq.x;
                    const float d = length(D);
                    o = gq.x;
                    dir = D / d;
                    tmax = d - kShadowEps;
                    valid = true;
                    zm = Ms[k];
                }
            }
            B.zm[(size_t)k * npix + pix] = zm;
        }
        emit_ray(B.list, &nray, valid, o, dir, tmax, (uint32_t)k * npix + pix);
    }
    publish_rays(B.list, &nray);
}

__global__ __launch_bounds__(kGiSpatialBlock) void gi_spatial_resolve_kernel(GiArgs A, GiSpatialBufs B) {
    const LaunchInfo &li = A.li;
    const TileOrder tiles(li.width, li.height);
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t npix = li.width * li.height;
    uint32_t px, py;
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;
    const uint32_t pix = py * li.width + px;
    const GPix gb = load_g(A.g, pix);
    const int RM = A.cur.M(pix);
    if (gb.valid && RM > 0) {
        GRes m = A.prev.load(pix);
        int Z = RM;
        for (int k = 0; k < A.neighbours; k++) {
            const size_t j = (size_t)k * npix + pix;
            const int zm = B.zm[j];
            if (zm > 0 && B.occ[j] == 0) Z += zm;
        }
        const float den = gi_target(m.s, gb.x, gb.n) * (float)Z;
        m.W = (den > 0.0f) ? m.w / den : 0.0f;
        m.xo = gb.x;
        A.prev.store(pix, m);
        if (A.shade) gi_shade(A, pix, gb, m);
    } else {
        A.prev.reset(pix);
        if (A.shade) {
            GRes z;
            z.W = 0.0f;
            gi_shade(A, pix, gb, z);
        }
    }
}

__global__ void gi_seed_kernel(GPlanes a, GPlanes b, uint32_t w, uint32_t h, uint32_t user) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    GRes r;
    r.s.xs = r.s.ns = r.s.Lo = mk(0, 0, 0);
    r.xo = mk(0, 0, 0);
    r.w = r.W = 0.0f;
    r.M = 0;
    r.seed = reservoir_seed(i % w, i / w, user ^ 0x47490000u);
    a.store(i, r);
    b.store(i, r);
}

__global__ void gi_unpack_kernel(GPlanes a, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GRes r = a.load(i);
    float v[20] = {r.s.xs.x, r.s.xs.y, r.s.xs.z, r.s.ns.x, r.s.ns.y, r.s.ns.z, r.s.Lo.x, r.s.Lo.y, r.s.Lo.z,
                   r.xo.x,   r.xo.y,   r.xo.z,   r.w,      r.W,      (float)r.M, r.seed.x, r.seed.y, r.seed.z,
                   0.0f,     0.0f};
    for (int k = 0; k < 20; k++) out[20 * (size_t)i + k] = v[k];
}

}  // namespace

void gi_alloc(Context &ctx, GiState &st, uint32_t w, uint32_t h, uint32_t user_seed) {
    size_t n = (size_t)w * h;
    st.w = w;
    st.h = h;
    for (int b = 0; b < 2; b++) {
        st.res[b].alloc(sizeof(float4) * 5 * n);
        st.gxd[b].alloc(sizeof(float4) * n);
        st.gnv[b].alloc(sizeof(float4) * n);
        st.gmat[b].alloc(sizeof(uint32_t) * n);
        st.guv[b].alloc(sizeof(float2) * n);
        KRT_HIP(hipMemsetAsync(st.guv[b].p, 0, st.guv[b].bytes, ctx.stream));
        KRT_HIP(hipMemsetAsync(st.gxd[b].p, 0, st.gxd[b].bytes, ctx.stream));
        KRT_HIP(hipMemsetAsync(st.gnv[b].p, 0, st.gnv[b].bytes, ctx.stream));
        KRT_HIP(hipMemsetAsync(st.gmat[b].p, 0, st.gmat[b].bytes, ctx.stream));
    }
    st.direct.alloc(sizeof(float4) * n);
    st.aux.alloc(sizeof(float4) * n);
    st.cur = 0;
    st.gcur = 0;
    GPlanes a{st.res[0].as<float4>(), (uint32_t)n}, b{st.res[1].as<float4>(), (uint32_t)n};
    gi_seed_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx.stream>>>(a, b, w, h, user_seed);
    KRT_HIP(hipGetLastError());
    for (int k = 0; k < 16; k++) st.prev_pv[k] = 0.0f;
}

void gi_read(Context &ctx, GiState &st, int which, float *host_out) {
    KRT_REQUIRE(st.w > 0, KRT_ERR_STATE, "ReSTIR GI buffers not loaded");
    uint32_t n = st.w * st.h;
    int b = which == 0 ? 1 - st.cur : st.cur;
    DevBuf tmp(sizeof(float) * 20 * (size_t)n);
    gi_unpack_kernel<<<(n + 255) / 256, 256, 0, ctx.stream>>>(GPlanes{st.res[b].as<float4>(), n}, n, tmp.as<float>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipMemcpyAsync(host_out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost, ctx.stream));
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

void gi_render(Context &ctx, Scene &s, const LaunchInfo &li_in, const RestirOptions &opt, GiState &st,
               unsigned long long *ctr, PassTimes &times) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "restir_gi renders flat scenes (krt_scene_upload)");
    KRT_REQUIRE(!li_in.env, KRT_ERR_STATE,
                "restir_gi has no environment lighting (its samples are surface points); disable the envmap");
    KRT_REQUIRE(st.w == li_in.width && st.h == li_in.height, KRT_ERR_STATE, "ReSTIR GI buffers not loaded");
    KRT_REQUIRE(opt.neighbours >= 0 && opt.neighbours <= kGiMaxNeighbours, KRT_ERR_INVALID,
                "restir_gi neighbours must be in [0, 8]");
    KRT_REQUIRE(opt.spatial_passes >= 0 && opt.m_cap >= 0 && opt.radius >= 0, KRT_ERR_INVALID,
                "restir_gi options must be non-negative");
    LaunchInfo li = li_in;
    for (int k = 0; k < 16; k++) li.prev_pv[k] = st.prev_pv[k];
    const uint32_t npix = li.width * li.height;
    uint32_t nb = persistent_blocks(npix, kTraceBlock, kMaxResidentBlocks);
    uint32_t nthreads = nb * kTraceBlock;
    auto planes = [&](int b) { return GPlanes{st.res[b].as<float4>(), npix}; };
    auto gbuf = [&](int b) {
        return GBuf{st.gxd[b].as<float4>(), st.gnv[b].as<float4>(), st.gmat[b].as<uint32_t>(), st.guv[b].as<float2>()};
    };
    GiArgs A;
    A.li = li;
    A.sv = make_view(s);
    A.g = gbuf(st.gcur);
    A.pg = gbuf(1 - st.gcur);
    A.direct = st.direct.as<float4>();
    A.aux = st.aux.as<float4>();
    // one spill slab for the temporal kernel (32-entry LDS stacks) and the
    // list walks (16 entries, up to kMaxResidentBlocks workgroups): size it
    // for both first
    const uint32_t items = ((li.width + 7u) / 8u) * ((li.height + 7u) / 8u) * 64u;
    const uint32_t gb = (items + kGiSpatialBlock - 1) / kGiSpatialBlock;  // spatial producers = list segments
    spill_buffer(ctx, s, 2 * kMaxResidentBlocks * kTraceBlock, kLdsDepthLean);
    A.spill = spill_buffer(ctx, s, nthreads);
    A.ctr = ctr;
    A.m_cap = opt.m_cap;
    A.neighbours = opt.neighbours;
    A.radius = opt.radius;
    hipEvent_t ev[4];
    for (int k = 0; k < 4; k++) KRT_HIP(hipEventCreate(&ev[k]));
    KRT_HIP(hipEventRecord(ev[0], ctx.stream));
    A.cur = planes(st.cur);
    A.prev = planes(1 - st.cur);
    A.shade = 0;
    {
        WaveScratch &ws = ctx.wave;
        const size_t levels = (size_t)li.max_depth + 1;
        const size_t cap = (size_t)gb * kGiWaveBlock;
        auto grow = [](DevBuf &b, size_t bytes) {
            if (b.bytes < bytes) b.alloc(bytes);
        };
        grow(ws.hits, sizeof(float4) * npix);
        grow(ws.state, sizeof(float4) * npix);
        grow(ws.hist, sizeof(float4) * 3 * levels * npix);
        grow(ws.occ, levels * npix);
        grow(ws.v0, sizeof(float4) * (s.textured ? 5 : 4) * npix);
        for (int l = 0; l < 2; l++) {
            grow(ws.rays[l], sizeof(float4) * 2 * cap);
            grow(ws.counts[l], sizeof(uint32_t) * gb);
            grow(ws.prefix[l], sizeof(uint32_t) * (gb + 1));
            grow(ws.dense[l], sizeof(float4) * 2 * cap);
        }
        GiWave W;
        W.cl = RayList{ws.rays[0].as<float4>(), ws.counts[0].as<uint32_t>(), (uint32_t)kGiWaveBlock, gb,
                       ws.prefix[0].as<uint32_t>(), ws.dense[0].as<float4>()};
        W.sl = RayList{ws.rays[1].as<float4>(), ws.counts[1].as<uint32_t>(), (uint32_t)kGiWaveBlock, gb,
                       ws.prefix[1].as<uint32_t>(), ws.dense[1].as<float4>()};
        W.hits = ws.hits.as<float4>();
        W.state = ws.state.as<float4>();
        W.hist = ws.hist.as<float4>();
        W.v0 = ws.v0.as<float4>();
        W.occ = ws.occ.as<uint8_t>();
        gi_camera_kernel<<<gb, kGiWaveBlock, 0, ctx.stream>>>(li, W);
        KRT_HIP(hipGetLastError());
        walk_closest_list(ctx, false, A.sv, W.cl, W.hits, nullptr, A.spill, ctr, 0);
        for (int level = 0; level <= li.max_depth; level++) {
            gi_shade_kernel<<<gb, kGiWaveBlock, 0, ctx.stream>>>(li, A.sv, W, level);
            KRT_HIP(hipGetLastError());
            // the level's shadow rays and the next level's closest rays
            if (level < li.max_depth)
                walk_pair(ctx, false, A.sv, W.cl, W.hits, nullptr, W.sl, W.occ + (size_t)level * npix, A.spill, ctr,
                          0);
            else
                walk_shadow_list(ctx, false, A.sv, W.sl, W.occ + (size_t)level * npix, A.spill, ctr, 0);
        }
        gi_compose_kernel<<<gb, kGiWaveBlock, 0, ctx.stream>>>(A, W);
    }
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(ev[1], ctx.stream));
    A.shade = opt.spatial_passes == 0;
    gi_temporal_kernel<<<nb, kTraceBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipEventRecord(ev[2], ctx.stream));
    if (opt.spatial_passes > 0) {
        const int S = opt.neighbours;
        const size_t list_slots = (size_t)gb * kGiSpatialBlock * (size_t)S + 1;
        const size_t slots = (size_t)S * npix + 1;
        if (st.rays.bytes < sizeof(float4) * 2 * list_slots) st.rays.alloc(sizeof(float4) * 2 * list_slots);
        if (st.seg_counts.bytes < sizeof(uint32_t) * gb) st.seg_counts.alloc(sizeof(uint32_t) * gb);
        if (st.seg_prefix.bytes < sizeof(uint32_t) * (gb + 1)) st.seg_prefix.alloc(sizeof(uint32_t) * (gb + 1));
        if (st.dense.bytes < sizeof(float4) * 2 * list_slots) st.dense.alloc(sizeof(float4) * 2 * list_slots);
        if (st.occ.bytes < slots) st.occ.alloc(slots);
        if (st.zm.bytes < sizeof(int32_t) * slots) st.zm.alloc(sizeof(int32_t) * slots);
        GiSpatialBufs B;
        B.list.rays = st.rays.as<float4>();
        B.list.counts = st.seg_counts.as<uint32_t>();
        B.list.prefix = st.seg_prefix.as<uint32_t>();
        B.list.dense = st.dense.as<float4>();
        B.list.seg = kGiSpatialBlock * (uint32_t)S;
        B.list.nseg = gb;
        B.occ = st.occ.as<uint8_t>();
        B.zm = st.zm.as<int32_t>();
        for (int pass = 0; pass < opt.spatial_passes; pass++) {
            A.cur = planes(st.cur);
            A.prev = planes(1 - st.cur);
            A.shade = pass == opt.spatial_passes - 1;
            gi_spatial_rays_kernel<<<gb, kGiSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            walk_shadow_list(ctx, false, A.sv, B.list, B.occ, A.spill, ctr, 2);
            gi_spatial_merge_kernel<<<gb, kGiSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            walk_shadow_list(ctx, false, A.sv, B.list, B.occ, A.spill, ctr, 2);
            gi_spatial_resolve_kernel<<<gb, kGiSpatialBlock, 0, ctx.stream>>>(A, B);
            KRT_HIP(hipGetLastError());
            st.cur = 1 - st.cur;
        }
    }
    KRT_HIP(hipEventRecord(ev[3], ctx.stream));
    st.cur = 1 - st.cur;
    st.gcur = 1 - st.gcur;
    for (int k = 0; k < 16; k++) st.prev_pv[k] = li.pv[k];
    KRT_HIP(hipEventSynchronize(ev[3]));
    for (int k = 0; k < 3; k++) KRT_HIP(hipEventElapsedTime(&times.ms[k], ev[k], ev[k + 1]));
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
}

}  // namespace krt

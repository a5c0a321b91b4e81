// Per-hit shading shared by the path-tracer and ReSTIR kernels:
// surface reconstruction (LOAD_INTERSECTION_DATA, source/optix/amadeus_common.cuh:
// 490-515 with calculate_normal :74-112 and calculate_intersection :62-71, for
// world-space triangles), next-event estimation Ld (:255-286 with Ld_light,
// include/optix/lighting.cuh:149-174) and the camera ray (make_ray,
// source/optix/amadeus_path_tracer.cu:26-56).
#pragma once

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_traverse.h"

namespace krt {

constexpr float kEps = 1e-3f;        // amadeus_common.cuh:22
constexpr float kShadowEps = 0.05f;  // amadeus_common.cuh:193
constexpr float kTraceTmax = 1e16f;  // amadeus_common.cuh:472

struct RayCount {
    unsigned closest = 0, shadow = 0;
};

struct Surface {
    f3 x, n;
    bool entering;
    uint32_t mat_id;
};

KRT_DEV f3 interp(f3 a, f3 b, f3 c, float b1, float b2) {
    float w = 1.0f - b1 - b2;
    return (w * a + b1 * b) + b2 * c;
}

// Instanced hit: calculate_intersection / calculate_normal with the
// instance's model matrix (amadeus_common.cuh:62-112): x = M (interp, 1), the
// edges and the interpolated normal through M with w = 0.
KRT_DEV bool load_surface_inst(const SceneView &sv, const Hit &h, f3 dir, Surface &sf) {
    const float4 *rec = sv.inst + kInstRec * (size_t)h.inst;
    const float4 m0 = rec[3], m1 = rec[4], m2 = rec[5], meta = rec[6], geo = rec[7];
    const uint32_t local = (uint32_t)(h.prim - __float_as_int(meta.z));
    const uint32_t *ix = sv.idx + 3 * ((size_t)__float_as_int(geo.y) + local);
    const uint32_t vb = (uint32_t)__float_as_int(geo.x);
    const uint32_t i0 = vb + ix[0], i1 = vb + ix[1], i2 = vb + ix[2];
    f3 a = ld3(sv.pos + 3 * (size_t)i0), b = ld3(sv.pos + 3 * (size_t)i1), c = ld3(sv.pos + 3 * (size_t)i2);
    f3 ng = normalize(cross(xform(m0, m1, m2, b - a, 0.0f), xform(m0, m1, m2, c - a, 0.0f)));
    if (dot(ng, dir) > 0.0f) {
        ng = -ng;
        sf.entering = false;
    } else {
        sf.entering = true;
    }
    ng = normalize(ng);
    f3 nn = xform(m0, m1, m2,
                  interp(ld3(sv.nrm + 3 * (size_t)i0), ld3(sv.nrm + 3 * (size_t)i1), ld3(sv.nrm + 3 * (size_t)i2),
                         h.b1, h.b2),
                  0.0f);
    if (dot(nn, ng) < 0.0f) nn = -nn;
    sf.n = normalize(nn);
    sf.x = xform(m0, m1, m2, interp(a, b, c, h.b1, h.b2), 1.0f);
    sf.mat_id = sv.tri_mat[h.prim];
    return !isnan3(sf.n);
}

// Returns false when the shading normal is NaN (the reference then returns 0).
template <bool INST = false>
KRT_DEV bool load_surface(const SceneView &sv, const Hit &h, f3 dir, Surface &sf) {
    if (INST) return load_surface_inst(sv, h, dir, sf);
    const uint32_t i0 = sv.idx[3 * h.prim], i1 = sv.idx[3 * h.prim + 1], i2 = sv.idx[3 * h.prim + 2];
    f3 a = ld3(sv.pos + 3 * i0), b = ld3(sv.pos + 3 * i1), c = ld3(sv.pos + 3 * i2);
    f3 ng = normalize(cross(b - a, c - a));
    if (dot(ng, dir) > 0.0f) {
        ng = -ng;
        sf.entering = false;
    } else {
        sf.entering = true;
    }
    ng = normalize(ng);
    f3 nn = interp(ld3(sv.nrm + 3 * i0), ld3(sv.nrm + 3 * i1), ld3(sv.nrm + 3 * i2), h.b1, h.b2);
    if (dot(nn, ng) < 0.0f) nn = -nn;
    sf.n = normalize(nn);
    sf.x = interp(a, b, c, h.b1, h.b2);
    sf.mat_id = sv.tri_mat[h.prim];
    return !isnan3(sf.n);
}

KRT_DEV f3 offset_point(f3 x, f3 n, int type) {
    return x + ((type == eTransmission ? -1.0f : 1.0f) * n) * kEps;
}

// Closest hit of the scene: one-level walk of the binary tree, or the
// two-level walk of an instanced scene.
template <bool INST = false>
KRT_DEV bool trace_closest(const SceneView &sv, f3 o, f3 d, float tmax, const TraceStack &stk, Hit &h,
                           TravStats &ts) {
    if (INST) return traverse_inst<false>(sv.tlas, sv.inst, sv.wide, sv.tri, o, d, tmax, stk, h, ts);
    return traverse<false>(sv.nodes, sv.wide, sv.tri, o, d, tmax, stk, h, ts);
}

template <bool WIDE = true, bool INST = false>
KRT_DEV bool occluded(const SceneView &sv, f3 o, f3 d, float tmax, const TraceStack &stk, TravStats &ts) {
    Hit h;
    if (INST) return traverse_inst<true>(sv.tlas, sv.inst, sv.wide, sv.tri, o, d, tmax, stk, h, ts);
    return traverse<true, WIDE>(sv.nodes, sv.wide, sv.tri, o, d, tmax, stk, h, ts);
}

// The environment map of a launch (ArmadaLaunchInfo::environment_map,
// include/amadeus/armada.cuh:76-77); tex == nullptr when it is off.
KRT_DEV EnvMap env_of(const LaunchInfo &li) { return EnvMap{li.env, li.env_w, li.env_h}; }

// Ld's light sample with its shadow ray left to the caller (Ld,
// amadeus_common.cuh:255-286): one light of the triangle lights and the
// environment map chosen uniformly; a triangle light through Ld_light
// (lighting.cuh:149-174), the environment through Ld_Environment
// (amadeus_common.cuh:215-252: a direction about world +z, pdf 1/(4 pi), the
// shadow ray to WORLD_RADIUS).  contr is the unshadowed contribution, pdf the
// probability it is divided by; false when there is no light at all.
struct LdSample {
    f3 contr, wi;
    float pdf, tmax;
};
KRT_DEV bool ld_sample(const SceneView &sv, const EnvMap &env, const SurfaceHit &sh, f3 &seed, LdSample &s) {
    const unsigned total = sv.nl + (env.tex ? 1u : 0u);
    if (total == 0) {
        rand_uniform(seed);
        return false;
    }
    const unsigned i = pick_light(seed, total);
    float pdf = 1.0f / (float)total;
    if (i < sv.nl) {
        TriLight L = load_light(sv.lights, i);
        float light_pdf = 1.0f / L.area;
        f3 lpos = sample_area_light(L, seed);
        f3 wi = normalize(lpos - sh.x);
        float R = length(lpos - sh.x);
        f3 f = brdf(sh, wi, eDiffuse);
        float ldot = fabsf(dot(L.normal, wi));
        float geometric = ldot * fabsf(dot(sh.n, wi)) / (R * R);
        s.contr = f * L.intensity * geometric;
        s.wi = wi;
        s.tmax = R - kShadowEps;
        pdf *= light_pdf;
    } else {
        const f3 wi = env_sample_dir(seed);
        const f3 Li = env_radiance(env, wi, false);
        const f3 f = brdf(sh, wi, eDiffuse) * fabsf(dot(sh.n, wi));
        s.contr = f * Li;
        s.wi = wi;
        s.tmax = kEnvDistance - kShadowEps;
        pdf *= kEnvPdf;
    }
    s.pdf = pdf;
    return true;
}

// Ld with its shadow ray traced here (the fused kernels).
template <bool WIDE = true, bool INST = false>
KRT_DEV f3 Ld(const SceneView &sv, const EnvMap &env, const SurfaceHit &sh, f3 &seed, const TraceStack &stk,
              TravStats &ts, RayCount &rc) {
    LdSample s;
    if (!ld_sample(sv, env, sh, seed, s)) return mk(0.0f, 0.0f, 0.0f);
    rc.shadow++;
    const f3 contr = occluded<WIDE, INST>(sv, sh.x, s.wi, s.tmax, stk, ts) ? mk(0.0f, 0.0f, 0.0f) : s.contr;
    return contr / s.pdf;
}

// Camera ray; the seed is advanced once by pcg3f (jittered or fixed -0.5 offset).
KRT_DEV void make_ray(const LaunchInfo &li, uint32_t px, uint32_t py, f3 &seed, bool jitter, f3 &o, f3 &d) {
    pcg3f(seed);
    float xoff = jitter ? fract(seed.x) - 0.5f : -0.5f;
    float yoff = jitter ? fract(seed.y) - 0.5f : -0.5f;
    float dx = 2.0f * (((float)px + xoff) / (float)li.width) - 1.0f;
    float dy = 2.0f * (((float)py + yoff) / (float)li.height) - 1.0f;
    o = ld3(li.center);
    d = normalize((dx * ld3(li.U) + dy * ld3(li.V)) + ld3(li.W));
}

// sample_spatial_neighborhood (source/optix/amadeus_restir.cu:295-314): a
// point in a disk of `radius` pixels, clamped to the image.
KRT_DEV uint32_t spatial_neighbour_px(const LaunchInfo &li, uint32_t index, f3 &seed, int radius) {
    int width = (int)li.width, height = (int)li.height;
    int x = (int)(index % li.width), y = (int)(index / li.width);
    f3 off = rand_uniform_3f(seed);
    float c, s;
    sincos_turn(off.x, c, s);
    float rr = (float)radius * sqrtf(off.y);
    int x0 = (int)((float)x + rr * c);
    int y0 = (int)((float)y + rr * s);
    x0 = min(max(x0, 0), width - 1);
    y0 = min(max(y0, 0), height - 1);
    return (uint32_t)(x0 + y0 * width);
}

// Path history of one lane: per level {direct, has_pdf} and {T}.  value is
// then evaluated from the deepest level up, value_k = direct_k + T_k * value_k+1
// exactly as the reference's recursion returns it.
struct History {
    float4 *base;
    uint32_t stride;  // lanes in the grid
    KRT_DEV void put(int level, f3 direct, bool has_pdf, f3 T) const {
        base[(2 * level) * stride] = make_float4(direct.x, direct.y, direct.z, has_pdf ? 1.0f : 0.0f);
        base[(2 * level + 1) * stride] = make_float4(T.x, T.y, T.z, 0.0f);
    }
    KRT_DEV void get(int level, f3 &direct, bool &has_pdf, f3 &T) const {
        float4 a = base[(2 * level) * stride], b = base[(2 * level + 1) * stride];
        direct = mk(a.x, a.y, a.z);
        has_pdf = a.w != 0.0f;
        T = mk(b.x, b.y, b.z);
    }
};

}  // namespace krt

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

// Material textures (calculate_material / calculate_normal,
// source/optix/amadeus_common.cuh:74-187): RGBA8 texels (r in the low byte)
// read as byte / 255, with tex2D <float4>'s normalised coordinates, bilinear
// filtering and wrap addressing (the descriptor of import_vulkan_texture,
// include/cuda/interop.cuh:11-51), filtered in FP32 (the texture unit's 8-bit
// weights are not restated).  A texture is {first texel, width, height} in
// one texel array (TexSet); coordinates beyond +-2^16 (or NaN) read zero.
KRT_DEV float4 tex_fetch(const TexSet &ts, uint32_t id, float u, float v) {
    if (!(fabsf(u) <= 65536.0f && fabsf(v) <= 65536.0f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint4 d = ts.desc[id];
    const uint32_t *tx = ts.texels + d.x;
    const float x = u * (float)d.y - 0.5f, y = v * (float)d.z - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int ix = (int)fx, iy = (int)fy, W = (int)d.y, H = (int)d.z;
    int i0 = ix % W, i1 = (ix + 1) % W, j0 = iy % H, j1 = (iy + 1) % H;
    i0 += i0 < 0 ? W : 0;
    i1 += i1 < 0 ? W : 0;
    j0 += j0 < 0 ? H : 0;
    j1 += j1 < 0 ? H : 0;
    const float4 t00 = texel8(tx[(size_t)j0 * d.y + i0]), t10 = texel8(tx[(size_t)j0 * d.y + i1]);
    const float4 t01 = texel8(tx[(size_t)j1 * d.y + i0]), t11 = texel8(tx[(size_t)j1 * d.y + i1]);
    const float a0 = 1.0f - a, b0 = 1.0f - b;
    const float w00 = a0 * b0, w10 = a * b0, w01 = a0 * b, w11 = a * b;
    return make_float4((t00.x * w00 + t10.x * w10) + (t01.x * w01 + t11.x * w11),
                       (t00.y * w00 + t10.y * w10) + (t01.y * w01 + t11.y * w11),
                       (t00.z * w00 + t10.z * w10) + (t01.z * w01 + t11.z * w11),
                       (t00.w * w00 + t10.w * w10) + (t01.w * w01 + t11.w * w11));
}

// calculate_material's texture reads (amadeus_common.cuh:154-187): diffuse,
// specular and emission take the texel's rgb, roughness its r.
KRT_DEV void apply_textures(const TexSet &ts, uint4 t, float u, float v, Material &M) {
    const uint32_t dif = t.x & 0xFFFFu, spe = t.y & 0xFFFFu, emi = t.y >> 16, rou = t.z & 0xFFFFu;
    if (dif != kNoTex) {
        const float4 c = tex_fetch(ts, dif, u, v);
        M.diffuse = mk(c.x, c.y, c.z);
    }
    if (spe != kNoTex) {
        const float4 c = tex_fetch(ts, spe, u, v);
        M.specular = mk(c.x, c.y, c.z);
    }
    if (emi != kNoTex) {
        const float4 c = tex_fetch(ts, emi, u, v);
        M.emission = mk(c.x, c.y, c.z);
    }
    if (rou != kNoTex) M.roughness = tex_fetch(ts, rou, u, v).x;
}

struct Surface {
    f3 x, n;
    bool entering;
    uint32_t mat_id;
    float u, v;  // texture coordinates (calculate_material's uv; 0 without textures)
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
    sf.u = sf.v = 0.0f;
    return !isnan3(sf.n);
}

// The texture coordinates of a hit (LOAD_INTERSECTION_DATA,
// amadeus_common.cuh:498-502: interpolated, v flipped) and calculate_normal's
// normal map (:114-146): n = 2 rgb - 1 in the frame of the interpolated
// tangent and bitangent (each normalised twice, as the reference does; the
// model matrix is the identity for world-space triangles) and the shading
// normal.
KRT_DEV void surface_uv(const SceneView &sv, uint32_t i0, uint32_t i1, uint32_t i2, const Hit &h, Surface &sf,
                        bool normal_map) {
    const float2 ta = sv.uv[i0], tb = sv.uv[i1], tc = sv.uv[i2];
    const float w = 1.0f - h.b1 - h.b2;
    sf.u = (w * ta.x + h.b1 * tb.x) + h.b2 * tc.x;
    sf.v = 1.0f - ((w * ta.y + h.b1 * tb.y) + h.b2 * tc.y);
    if (!normal_map) return;
    const uint32_t nmap = material_tex(sv.mats, sf.mat_id).x >> 16;
    if (nmap == kNoTex || !sv.tan) return;
    const float4 t4 = tex_fetch(sv.tex, nmap, sf.u, sf.v);
    const f3 tn = mk(2.0f * t4.x - 1.0f, 2.0f * t4.y - 1.0f, 2.0f * t4.z - 1.0f);
    const float *ta6 = sv.tan + 6 * (size_t)i0, *tb6 = sv.tan + 6 * (size_t)i1, *tc6 = sv.tan + 6 * (size_t)i2;
    const f3 T = normalize(normalize(interp(ld3(ta6), ld3(tb6), ld3(tc6), h.b1, h.b2)));
    const f3 B = normalize(normalize(interp(ld3(ta6 + 3), ld3(tb6 + 3), ld3(tc6 + 3), h.b1, h.b2)));
    const f3 N = normalize(sf.n);
    sf.n = normalize((T * tn.x + B * tn.y) + N * tn.z);
}

// Returns false when the shading normal is NaN (the reference then returns 0).
// NMAP = false: the editor's ray-cast G-buffer (editor/gbuffer_rtx_shader.cu:
// 95-160), which records (u, 1 - v) but applies no normal map.
template <bool INST = false, bool NMAP = true>
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
    sf.u = sf.v = 0.0f;
    if (sv.uv) surface_uv(sv, i0, i1, i2, h, sf, NMAP);
    return !isnan3(sf.n);
}

// calculate_material (amadeus_common.cuh:154-187) of material m at (u, v).
KRT_DEV Material textured_material(const SceneView &sv, uint32_t m, float u, float v) {
    Material M = load_material(sv.mats, m);
    const uint4 t = material_tex(sv.mats, m);
    if (t.w) apply_textures(sv.tex, t, u, v, M);
    return M;
}
KRT_DEV Material surface_material(const SceneView &sv, const Surface &sf) {
    return textured_material(sv, sf.mat_id, sf.u, sf.v);
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

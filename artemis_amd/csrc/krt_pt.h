// One path-tracer sample of one pixel: the estimator of
// source/optix/amadeus_path_tracer.cu:73-228 (raygen + closest-hit + miss),
// run by the batched serving kernel (krt_serve.hip).  The attachment's split
// passes (krt_pt.hip) evaluate the same expressions in the same order, so a
// pixel's value does not depend on which one ran it.
//
//   value_k = Le_k + Ld_k (+ T_k * value_{k+1} when pdf_k > 0),
//   T_k = f_k |wi.n| / pdf_k,
// with shading at depths 0..max_depth; the ray past max_depth contributes 0 and
// is not traced (its program only writes value = 0, LOAD_RAYPACKET).  Per-level
// (Le+Ld, T) go to a per-lane history slab and the value is folded from the
// deepest level up, so the sum is rounded in the reference's recursive order.
#pragma once

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {

// `sample` is the frame's sample index (li.samples for one frame per launch;
// the serving kernel runs several consecutive frames of a pixel in one thread,
// which leaves every accumulation step as a launch per frame would make it).
KRT_DEV void pt_pixel(const LaunchInfo &li, const SceneView &sv, uint32_t pix, int sample, const TraceStack &stk,
                      const History &hist, TravStats &ts, RayCount &rc, double *acc64) {
    const float count = (float)sample;
    const uint32_t px = pix % li.width, py = pix / li.width;
    f3 seed = pixel_seed(px, py, (uint32_t)sample, li.user_seed);
    f3 o, d;
    make_ray(li, px, py, seed, li.jitter != 0, o, d);
    f3 pos0 = mk(0, 0, 0), n0 = mk(0, 0, 0), alb0 = mk(0, 0, 0);
    const EnvMap env = env_of(li);
    int nlev = 0;
    for (int level = 0; level <= li.max_depth; level++) {
        Hit h;
        rc.closest++;
        traverse<false>(sv.nodes, sv.wide, sv.tri, o, d, kTraceTmax, stk, h, ts);
        if (h.prim < 0) {  // __miss__: the environment's radiance, or 0
            if (env.tex) {
                hist.put(level, env_radiance(env, d, true), false, mk(0.0f, 0.0f, 0.0f));
                nlev = level + 1;
            }
            break;
        }
        Surface sf;
        if (!load_surface(sv, h, d, sf)) break;
        Material mat = surface_material(sv, sf);
        f3 n = sf.n;
        f3 x = offset_point(sf.x, n, mat.type);
        SurfaceHit sh{mat, sf.entering, n, -d, x};
        f3 direct = mat.emission + Ld(sv, env, sh, seed, stk, ts, rc);
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
    else if (acc64) {  // converged-reference accumulator (FP64 only here)
        acc64[3 * (size_t)pix] += (double)value.x;
        acc64[3 * (size_t)pix + 1] += (double)value.y;
        acc64[3 * (size_t)pix + 2] += (double)value.z;
    }
    const bool acc = li.accumulate != 0;
    li.color[pix] = accumulate4(li.color[pix], col, acc, count);
    li.normal[pix] = accumulate4(li.normal[pix], make_float4(n0.x, n0.y, n0.z, 0.0f), acc, count);
    li.albedo[pix] = accumulate4(li.albedo[pix], make_float4(alb0.x, alb0.y, alb0.z, 0.0f), acc, count);
    li.position[pix] = make_float4(pos0.x, pos0.y, pos0.z, 0.0f);
}

}  // namespace krt

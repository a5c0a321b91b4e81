// The editor's G-buffer ReSTIR DI ("Mamba", editor/optix/mamba_shader.cu:35-505
// with the host sequence of editor/mamba.cu:1055-1150) as the attachment kind
// "mamba": a second ReSTIR semantics beside the Armada one (krt_restir.hip).
//
//   gbuffer   the surface of each pixel-centre camera ray (the editor rasters
//             it; the G-buffer texels hold position, normal with the
//             entering sign, triangle and material)
//   initial   (__raygen__direct_primary, :219-266) 8 candidates from
//             sample_light (an area light uniformly, one of its triangles
//             uniformly, pdf 1 / (triangle area x lights) -- the reference
//             leaves the triangle choice out of the pdf), weights target / pdf,
//             one visibility ray in the resampling target
//   temporal  (__raygen__temporal_reuse, :333-394) the previous frame's
//             reservoir at the reprojected pixel (when the camera moved),
//             capped at M = 200, occluded targets
//   spatial   (__raygen__spatial_reuse, :435-505) 2 neighbours in a 20 px
//             disk, capped at M = 200, 1/Z from the neighbours' surfaces
// The last enabled pass shades the pixel with one more ray (epsilon 1e-3,
// :38-63) and accumulates (Le s + emission + direct) / (s + 1) into the colour
// buffer.  One lane per pixel in 8x8 tiles; the few rays per pixel are walked
// in the kernels (any hit on the 4-wide tree).
//
// Deterministic where the reference races: the spatial pass reads the
// reservoirs the temporal pass left and writes a second buffer (swapped after
// the pass); the previous-frame reservoirs are double-buffered; the M cap of a
// reprojected reservoir applies to the reader's copy.  Seeds (x, y, time)
// take time = frame number + 4096 x user seed (the reference reads the wall
// clock); reservoir seeds start from reservoir_seed().  The CPU oracle
// (orc_mamba_*) restates the same.
#include <hip/hip_runtime.h>

#include <cstring>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_shade.h"
#include "krt_traverse.h"

namespace krt {

namespace {

constexpr int kMambaBlock = 128;
constexpr float kMambaEps = 1e-3f;
constexpr int kMambaCandidates = 8;
constexpr int kMambaMCap = 200;
constexpr int kMambaNeighbours = 2;
constexpr int kMambaRadius = 20;

struct MLight {  // LightInfo (mamba_shader.cuh:19-25)
    f3 position, normal, emission;
    float area;
};

struct MRes {  // Reservoir<LightInfo> (include/amadeus/reservoir.cuh:19-76)
    MLight data;
    float w, W;
    int M;
    f3 seed;
    KRT_DEV bool update(const MLight &s, float weight) {
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
    KRT_DEV void reset() {
        w = 0.0f;
        M = 0;
        W = 0.0f;
    }
};

// One 64 B record per pixel: {position, area}, {normal, w}, {emission, W}, {seed, M}
struct MBuf {
    float4 *p;
    KRT_DEV MRes load(uint32_t i) const {
        const float4 a = p[4 * (size_t)i], b = p[4 * (size_t)i + 1], c = p[4 * (size_t)i + 2],
                     d = p[4 * (size_t)i + 3];
        MRes r;
        r.data.position = mk(a.x, a.y, a.z);
        r.data.area = a.w;
        r.data.normal = mk(b.x, b.y, b.z);
        r.w = b.w;
        r.data.emission = mk(c.x, c.y, c.z);
        r.W = c.w;
        r.seed = mk(d.x, d.y, d.z);
        r.M = __float_as_int(d.w);
        return r;
    }
    KRT_DEV void store(uint32_t i, const MRes &r) const {
        p[4 * (size_t)i] = make_float4(r.data.position.x, r.data.position.y, r.data.position.z, r.data.area);
        p[4 * (size_t)i + 1] = make_float4(r.data.normal.x, r.data.normal.y, r.data.normal.z, r.w);
        p[4 * (size_t)i + 2] = make_float4(r.data.emission.x, r.data.emission.y, r.data.emission.z, r.W);
        p[4 * (size_t)i + 3] = make_float4(r.seed.x, r.seed.y, r.seed.z, __int_as_float(r.M));
    }
    KRT_DEV int M(uint32_t i) const { return __float_as_int(p[4 * (size_t)i + 3].w); }
};

struct MambaArgs {
    LaunchInfo li;
    SceneView sv;
    const int2 *ml_range;
    const int *ml_tris;
    const float4 *ml_emission;
    uint32_t nml;
    float4 *gx, *gn;  // G-buffer: {position, entering sign}, {normal, 0}
    int2 *gpm;        // {triangle, material}; material -1: no surface
    float2 *guv;      // texture coordinates (u, 1 - v)
    MBuf R, S, Pin, Pout;
    float time;
    int dirty, end_initial, end_temporal;
    int *spill;
};

KRT_DEV float lum(f3 v) { return v.x + v.y + v.z; }
KRT_DEV f3 cleanse(f3 v) { return isnan3(v) ? mk(0.0f, 0.0f, 0.0f) : v; }

KRT_DEV float sample_light(const MambaArgs &A, MLight &li, f3 &seed) {  // :99-162
    const unsigned total = A.nml;
    if (total == 0) return -1.0f;
    unsigned light_index = (unsigned)(rand_uniform(seed) * (float)total);
    light_index = light_index < total ? light_index : total - 1;
    const int2 rg = A.ml_range[light_index];
    const unsigned cnt = (unsigned)rg.y;
    unsigned ti = (unsigned)(rand_uniform(seed) * (float)cnt);
    ti = ti < cnt ? ti : cnt - 1;
    const int tri = A.ml_tris[rg.x + (int)ti];
    const uint32_t i0 = A.sv.idx[3 * tri], i1 = A.sv.idx[3 * tri + 1], i2 = A.sv.idx[3 * tri + 2];
    const f3 v0 = ld3(A.sv.pos + 3 * (size_t)i0), v1 = ld3(A.sv.pos + 3 * (size_t)i1), v2 = ld3(A.sv.pos + 3 * (size_t)i2);
    const f3 gnormal = normalize(cross(v1 - v0, v2 - v0));
    const f3 bary = pcg3f(seed);
    float u = bary.x, v = bary.y;
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    li.position = (v0 * (1.0f - u - v) + v1 * u) + v2 * v;
    li.normal = gnormal;
    li.emission = ld3(A.ml_emission[light_index]);
    li.area = 0.5f * length(cross(v1 - v0, v2 - v0));
    return 1.0f / (li.area * (float)total);
}

KRT_DEV f3 mamba_direct(const SurfaceHit &sh, const MLight &li) {  // direct_unoccluded (:70-83)
    const f3 wi = normalize(li.position - sh.x);
    const float R = length(li.position - sh.x);
    const f3 rho = cleanse(brdf(sh, wi, sh.mat.type));
    const float ldotn = fabsf(dot(li.normal, wi));
    const float ndotwi = fabsf(dot(sh.n, wi));
    const float falloff = 1.0f / (R * R);
    return (((rho * li.emission) * ldotn) * ndotwi) * falloff;
}

KRT_DEV bool mamba_occluded(const SceneView &sv, f3 origin, f3 target, const TraceStack &stk, TravStats &ts) {
    const f3 d = normalize(target - origin);
    return occluded<true, false>(sv, origin, d, length(target - origin) - kMambaEps, stk, ts);
}

KRT_DEV float occluded_target(const SceneView &sv, const SurfaceHit &sh, const MRes &r, const TraceStack &stk,
                              TravStats &ts) {
    const f3 Li = mamba_direct(sh, r.data);
    return lum(Li * (1.0f - (mamba_occluded(sv, sh.x, r.data.position, stk, ts) ? 1.0f : 0.0f)));
}

// load_surface_hit (:165-216) from the G-buffer
KRT_DEV bool mamba_surface(const MambaArgs &A, uint32_t i, SurfaceHit &sh) {
    const int2 pm = A.gpm[i];
    if (pm.y < 0) return false;
    const float4 gx = A.gx[i], gn = A.gn[i];
    const f3 origin = ld3(A.li.center), x = mk(gx.x, gx.y, gx.z);
    f3 n = mk(gn.x, gn.y, gn.z);
    if (dot(x - origin, n) > 0.0f) n = -n;
    sh.x = x;
    sh.wo = normalize(origin - x);
    sh.n = normalize(n);
    sh.entering = gx.w > 0.0f;
    sh.mat = load_material(A.sv.mats, (uint32_t)pm.y);
    // convert_material (editor/optix/common.cuh:34-47): the diffuse texture only
    const uint32_t dtex = material_tex(A.sv.mats, (uint32_t)pm.y).x & 0xFFFFu;
    if (A.sv.uv && dtex != kNoTex) {
        const float2 uv = A.guv[i];
        const float4 c = tex_fetch(A.sv.tex, dtex, uv.x, uv.y);
        sh.mat.diffuse = mk(c.x, c.y, c.z);
    }
    const float sign = (sh.mat.type == eTransmission) ? -1.0f : 1.0f;
    sh.x = sh.x + (sign * sh.n) * kMambaEps;
    return true;
}

KRT_DEV int mamba_reproject(const MambaArgs &A, int index, f3 pos) {  // :301-330
    if (!A.dirty) return index;
    const float *m = A.li.prev_pv;
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (m[0 + k] * pos.x + m[4 + k] * pos.y) + (m[8 + k] * pos.z + m[12 + k] * 1.0f);
    float u = v[0] / v[3], vv = v[1] / v[3];
    if (!(u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f)) return -1;
    u = (u + 1.0f) * 0.5f;
    vv = (vv + 1.0f) * 0.5f;
    int ix = round_half_up(u * (float)A.li.width);
    int iy = round_half_up(vv * (float)A.li.height);
    ix = min(max(ix, 0), (int)A.li.width - 1);
    iy = min(max(iy, 0), (int)A.li.height - 1);
    return iy * (int)A.li.width + ix;
}

KRT_DEV int mamba_neighbour(const LaunchInfo &li, int index, f3 &seed) {  // :411-432
    const int width = (int)li.width, height = (int)li.height;
    const int x = index % width, y = index / width;
    const f3 off = rand_uniform_3f(seed);
    float c, s;
    sincos_turn(off.x, c, s);
    const float rr = (float)kMambaRadius * sqrtf(off.y);
    int x0 = (int)((float)x + rr * c);
    int y0 = (int)((float)y + rr * s);
    x0 = min(max(x0, 0), width - 1);
    y0 = min(max(y0, 0), height - 1);
    return x0 + y0 * width;
}

KRT_DEV void mamba_shade(const MambaArgs &A, const SurfaceHit &sh, const MRes &r, uint32_t index,
                         const TraceStack &stk, TravStats &ts) {
    const f3 Ld = mamba_direct(sh, r.data);
    const bool occ = mamba_occluded(A.sv, sh.x, r.data.position, stk, ts);
    const f3 nd = cleanse((Ld * (1.0f - (occ ? 1.0f : 0.0f))) * r.W);
    const float s = (float)A.li.samples;
    const float4 c = A.li.color[index];
    const f3 v = ((mk(c.x, c.y, c.z) * s + sh.mat.emission) + nd) / (s + 1.0f);
    A.li.color[index] = make_float4(v.x, v.y, v.z, 1.0f);
}

#define MAMBA_PIXEL()                                                                              \
    __shared__ int lds_stack[kLdsDepth * kMambaBlock];                                             \
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;                                   \
    TraceStack stk{lds_stack + threadIdx.x, kMambaBlock, kLdsDepth, A.spill + gtid,                \
                   (int)(gridDim.x * blockDim.x)};                                                 \
    TravStats ts;                                                                                  \
    const TileOrder tiles(A.li.width, A.li.height);                                                \
    const uint32_t item = xcd_block() * blockDim.x + threadIdx.x;                                  \
    uint32_t px, py;                                                                               \
    if (item >= tiles.items || !tiles.pixel(item, px, py)) return;                                 \
    const uint32_t index = py * A.li.width + px;

__global__ __launch_bounds__(kMambaBlock) void mamba_gbuffer_kernel(MambaArgs A) {
    MAMBA_PIXEL();
    const LaunchInfo &li = A.li;
    const float dx = 2.0f * (((float)px + 0.5f) / (float)li.width) - 1.0f;
    const float dy = 2.0f * (((float)py + 0.5f) / (float)li.height) - 1.0f;
    const f3 o = ld3(li.center);
    const f3 d = normalize((dx * ld3(li.U) + dy * ld3(li.V)) + ld3(li.W));
    Hit h;
    trace_closest<false>(A.sv, o, d, kTraceTmax, stk, h, ts);
    Surface sf;
    float4 gx = make_float4(0.0f, 0.0f, 0.0f, 0.0f), gn = gx;
    int2 pm = make_int2(-1, -1);
    float2 uv = make_float2(0.0f, 0.0f);
    if (h.prim >= 0 && load_surface<false, false>(A.sv, h, d, sf)) {
        gx = make_float4(sf.x.x, sf.x.y, sf.x.z, sf.entering ? 1.0f : -1.0f);
        gn = make_float4(sf.n.x, sf.n.y, sf.n.z, 0.0f);
        pm = make_int2(h.prim, (int)sf.mat_id);
        uv = make_float2(sf.u, sf.v);
    }
    A.gx[index] = gx;
    A.gn[index] = gn;
    A.gpm[index] = pm;
    if (A.sv.uv) A.guv[index] = uv;
    const float4 c = A.li.color[index];
    A.li.color[index] = make_float4(c.x, c.y, c.z, 1.0f);
}

__global__ __launch_bounds__(kMambaBlock) void mamba_initial_kernel(MambaArgs A) {
    MAMBA_PIXEL();
    MRes r = A.R.load(index);
    r.reset();
    SurfaceHit sh;
    if (!mamba_surface(A, index, sh)) {
        A.R.store(index, r);
        return;
    }
    f3 seed = mk((float)px, (float)py, A.time);
    for (int i = 0; i < kMambaCandidates; i++) {
        MLight li;
        li.position = li.normal = li.emission = mk(0.0f, 0.0f, 0.0f);
        li.area = 0.0f;
        const float pdf = sample_light(A, li, seed);
        const f3 Ld = mamba_direct(sh, li);
        const float t = Ld.x + Ld.y + Ld.z;
        r.update(li, t / pdf > 0.0f ? t / pdf : 0.0f);
    }
    const f3 Ld = mamba_direct(sh, r.data);
    const bool occ = mamba_occluded(A.sv, sh.x, r.data.position, stk, ts);
    r.resample(lum(Ld * (1.0f - (occ ? 1.0f : 0.0f))));
    A.R.store(index, r);
    if (A.end_initial) mamba_shade(A, sh, r, index, stk, ts);
}

__global__ __launch_bounds__(kMambaBlock) void mamba_temporal_kernel(MambaArgs A) {
    MAMBA_PIXEL();
    const bool end = A.end_temporal != 0;
    SurfaceHit sh;
    if (!mamba_surface(A, index, sh)) {
        if (end) {
            MRes p = A.Pin.load(index);
            p.reset();
            A.Pout.store(index, p);
        }
        return;
    }
    const int pindex = mamba_reproject(A, (int)index, sh.x);
    MRes cur = A.R.load(index);
    if (cur.M <= 0) {
        if (end) A.Pout.store(index, A.Pin.load(index));
        return;
    }
    MRes merged;
    merged.data.position = merged.data.normal = merged.data.emission = mk(0.0f, 0.0f, 0.0f);
    merged.data.area = 0.0f;
    merged.w = merged.W = 0.0f;
    merged.M = 0;
    merged.seed = 2.0f * mk((float)px, (float)py, A.time);
    const float t_cur = occluded_target(A.sv, sh, cur, stk, ts);
    merged.update(cur.data, (t_cur * cur.W) * (float)cur.M);
    int N = cur.M;
    if (pindex >= 0 && A.Pin.M((uint32_t)pindex) > 0) {
        MRes prev = A.Pin.load((uint32_t)pindex);
        prev.M = min(prev.M, kMambaMCap);
        const float t_prev = occluded_target(A.sv, sh, prev, stk, ts);
        merged.update(prev.data, (t_prev * prev.W) * (float)prev.M);
        N += (t_prev > 0.0f ? 1 : 0) * prev.M;
    }
    merged.M = N;
    merged.resample(lum(mamba_direct(sh, merged.data)));
    A.R.store(index, merged);
    if (end) {
        mamba_shade(A, sh, merged, index, stk, ts);
        A.Pout.store(index, merged);
    }
}

__global__ __launch_bounds__(kMambaBlock) void mamba_spatial_kernel(MambaArgs A) {
    MAMBA_PIXEL();
    SurfaceHit sh;
    if (!mamba_surface(A, index, sh)) {
        MRes p = A.Pin.load(index);
        p.reset();
        A.Pout.store(index, p);
        A.S.store(index, A.R.load(index));
        return;
    }
    const MRes cur = A.R.load(index);
    if (cur.M <= 0) {  // the reference's debug colour
        A.li.color[index] = make_float4(0.0f, 0.0f, 1.0f, A.li.color[index].w);
        MRes p = A.Pin.load(index);
        p.reset();
        A.Pout.store(index, p);
        A.S.store(index, cur);
        return;
    }
    MRes merged;
    merged.data.position = merged.data.normal = merged.data.emission = mk(0.0f, 0.0f, 0.0f);
    merged.data.area = 0.0f;
    merged.w = merged.W = 0.0f;
    merged.M = 0;
    merged.seed = 2.0f * mk((float)px, (float)py, A.time);
    merged.update(cur.data, (lum(mamba_direct(sh, cur.data)) * cur.W) * (float)cur.M);
    SurfaceHit hits[kMambaNeighbours];
    int sizes[kMambaNeighbours];
    int count = 0;
    for (int i = 0; i < kMambaNeighbours; i++) {
        const int index0 = mamba_neighbour(A.li, (int)index, merged.seed);
        MRes nb = A.R.load((uint32_t)index0);
        nb.M = min(nb.M, kMambaMCap);
        const float t_nb = occluded_target(A.sv, sh, nb, stk, ts);
        if (nb.M > 0) merged.update(nb.data, (t_nb * nb.W) * (float)nb.M);
        hits[i].x = hits[i].n = hits[i].wo = mk(0.0f, 0.0f, 0.0f);
        mamba_surface(A, (uint32_t)index0, hits[i]);
        hits[i].wo = sh.wo;
        hits[i].entering = nb.M > 0;
        sizes[i] = nb.M;
        count += nb.M;
    }
    merged.M = cur.M + count;
    int Z = cur.M;
    for (int i = 0; i < kMambaNeighbours; i++) {
        const float t = hits[i].entering ? occluded_target(A.sv, hits[i], merged, stk, ts) : 0.0f;
        if (t > 0.0f) Z += sizes[i];
    }
    const float den = lum(mamba_direct(sh, merged.data)) * (float)Z;
    merged.W = den > 0.0f ? merged.w / den : 0.0f;
    A.S.store(index, merged);
    mamba_shade(A, sh, merged, index, stk, ts);
    A.Pout.store(index, merged);
}

__global__ void mamba_seed_kernel(MBuf a, MBuf b, MBuf c, MBuf d, uint32_t w, uint32_t h, uint32_t user) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    MRes r;
    r.data.position = r.data.normal = r.data.emission = mk(0.0f, 0.0f, 0.0f);
    r.data.area = 0.0f;
    r.w = r.W = 0.0f;
    r.M = 0;
    r.seed = reservoir_seed(i % w, i / w, user);
    a.store(i, r);
    b.store(i, r);
    c.store(i, r);
    d.store(i, r);
}

// {position, normal, emission, area, w, W, M, seed} per pixel
__global__ void mamba_unpack_kernel(MBuf a, uint32_t n, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MRes r = a.load(i);
    float *o = out + 16 * (size_t)i;
    o[0] = r.data.position.x, o[1] = r.data.position.y, o[2] = r.data.position.z;
    o[3] = r.data.normal.x, o[4] = r.data.normal.y, o[5] = r.data.normal.z;
    o[6] = r.data.emission.x, o[7] = r.data.emission.y, o[8] = r.data.emission.z;
    o[9] = r.data.area;
    o[10] = r.w, o[11] = r.W, o[12] = (float)r.M;
    o[13] = r.seed.x, o[14] = r.seed.y, o[15] = r.seed.z;
}

}  // namespace

void mamba_alloc(Context &ctx, MambaState &st, uint32_t w, uint32_t h, uint32_t user_seed) {
    const size_t n = (size_t)w * h;
    st = MambaState();
    st.w = w;
    st.h = h;
    st.user_seed = user_seed;
    for (auto *b : {&st.R, &st.S, &st.P[0], &st.P[1]}) b->alloc(sizeof(float4) * 4 * n);
    st.gx.alloc(sizeof(float4) * n);
    st.gn.alloc(sizeof(float4) * n);
    st.gpm.alloc(sizeof(int2) * n);
    st.guv.alloc(sizeof(float2) * n);
    mamba_seed_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx.stream>>>(
        MBuf{st.R.as<float4>()}, MBuf{st.S.as<float4>()}, MBuf{st.P[0].as<float4>()}, MBuf{st.P[1].as<float4>()}, w,
        h, user_seed);
    KRT_HIP(hipGetLastError());
}

void mamba_render(Context &ctx, Scene &s, const LaunchInfo &li_in, const MambaOptions &opt, MambaState &st,
                  PassTimes &times) {
    KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "mamba renders flat scenes (krt_scene_upload)");
    KRT_REQUIRE(st.w == li_in.width && st.h == li_in.height, KRT_ERR_STATE, "mamba buffers not loaded");
    MambaArgs A;
    A.li = li_in;
    // dirty (the camera moved since the last frame, mamba.cu:1056): reproject
    // and restart the accumulation (:1066-1068)
    float cam[12];
    std::memcpy(cam, li_in.center, sizeof(float) * 3);
    std::memcpy(cam + 3, li_in.U, sizeof(float) * 3);
    std::memcpy(cam + 6, li_in.V, sizeof(float) * 3);
    std::memcpy(cam + 9, li_in.W, sizeof(float) * 3);
    const bool dirty = st.have_prev && std::memcmp(cam, st.prev_cam, sizeof(cam)) != 0;
    for (int k = 0; k < 16; k++) A.li.prev_pv[k] = st.prev_pv[k];
    if (dirty || li_in.samples == 0) st.samples = 0;  // the armada restarts on !accumulate / a new scene
    A.li.samples = st.samples;
    A.sv = make_view(s);
    A.ml_range = s.ml_range.as<int2>();
    A.ml_tris = s.ml_tris.as<int>();
    A.ml_emission = s.ml_emission.as<float4>();
    A.nml = s.nml;
    A.gx = st.gx.as<float4>();
    A.gn = st.gn.as<float4>();
    A.gpm = st.gpm.as<int2>();
    A.guv = st.guv.as<float2>();
    A.R = MBuf{st.R.as<float4>()};
    A.S = MBuf{st.S.as<float4>()};
    A.Pin = MBuf{st.P[st.pin].as<float4>()};
    A.Pout = MBuf{st.P[1 - st.pin].as<float4>()};
    A.time = (float)st.frame + 4096.0f * (float)st.user_seed;
    A.dirty = dirty ? 1 : 0;
    A.end_initial = (!opt.temporal && !opt.spatial) ? 1 : 0;
    A.end_temporal = opt.spatial ? 0 : 1;
    const uint32_t items = ((li_in.width + 7u) / 8u) * ((li_in.height + 7u) / 8u) * 64u;
    const uint32_t gb = (items + kMambaBlock - 1) / kMambaBlock;
    A.spill = spill_buffer(ctx, s, gb * kMambaBlock);
    hipEvent_t e0, e1;
    KRT_HIP(hipEventCreate(&e0));
    KRT_HIP(hipEventCreate(&e1));
    KRT_HIP(hipEventRecord(e0, ctx.stream));
    mamba_gbuffer_kernel<<<gb, kMambaBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    mamba_initial_kernel<<<gb, kMambaBlock, 0, ctx.stream>>>(A);
    KRT_HIP(hipGetLastError());
    if (opt.temporal) {
        mamba_temporal_kernel<<<gb, kMambaBlock, 0, ctx.stream>>>(A);
        KRT_HIP(hipGetLastError());
    }
    if (opt.spatial) {
        mamba_spatial_kernel<<<gb, kMambaBlock, 0, ctx.stream>>>(A);
        KRT_HIP(hipGetLastError());
        std::swap(st.R, st.S);
    }
    if (opt.temporal || opt.spatial) st.pin = 1 - st.pin;
    KRT_HIP(hipEventRecord(e1, ctx.stream));
    KRT_HIP(hipEventSynchronize(e1));
    KRT_HIP(hipEventElapsedTime(&times.ms[0], e0, e1));
    times.frame = times.ms[0];
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < 16; k++) st.prev_pv[k] = li_in.pv[k];
    std::memcpy(st.prev_cam, cam, sizeof(cam));
    st.have_prev = true;
    st.samples++;
    st.frame++;
}

void mamba_read(Context &ctx, MambaState &st, int which, float *host_out) {
    KRT_REQUIRE(st.w > 0, KRT_ERR_STATE, "mamba buffers not loaded");
    const uint32_t n = st.w * st.h;
    DevBuf tmp(sizeof(float) * 16 * (size_t)n);
    const DevBuf &b = which == 0 ? st.R : st.P[st.pin];
    mamba_unpack_kernel<<<(n + 255) / 256, 256, 0, ctx.stream>>>(MBuf{(float4 *)b.p}, n, tmp.as<float>());
    KRT_HIP(hipGetLastError());
    KRT_HIP(hipMemcpyAsync(host_out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost, ctx.stream));
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

}  // namespace krt

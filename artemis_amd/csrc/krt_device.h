// Device math for the gfx950 kernels: vector helpers, pcg RNG, GGX/Glass BSDF,
// triangle lights, Möller–Trumbore.  Every expression keeps the evaluation order
// of the reference (include/cuda/math.cuh, random.cuh, brdf.cuh,
// include/optix/lighting.cuh); the library is compiled with -ffp-contract=off
// and IEEE division/sqrt, so results are bit-reproducible on any host that
// evaluates the same expressions in IEEE single precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KRT_DEV __device__ __forceinline__

namespace krt {

struct f3 {
    float x, y, z;
};

KRT_DEV f3 mk(float x, float y, float z) { return f3{x, y, z}; }
KRT_DEV f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
KRT_DEV f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
KRT_DEV f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
KRT_DEV f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
KRT_DEV f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
KRT_DEV f3 operator*(float s, f3 a) { return mk(a.x * s, a.y * s, a.z * s); }
KRT_DEV f3 operator/(f3 a, float s) {  // math.cuh:464-468: a * (1/s)
    float inv = 1.0f / s;
    return a * inv;
}
KRT_DEV f3 operator-(float a, f3 b) { return mk(a - b.x, a - b.y, a - b.z); }
KRT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KRT_DEV f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
KRT_DEV float length(f3 v) { return sqrtf(dot(v, v)); }
KRT_DEV f3 normalize(f3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
KRT_DEV f3 reflect(f3 i, f3 n) { return i - (2.0f * n) * dot(n, i); }
KRT_DEV float fract(float x) { return x - floorf(x); }
KRT_DEV float clampf(float x, float a, float b) {
    x = (x < a) ? a : x;
    return (x > b) ? b : x;
}
KRT_DEV float selmax(float a, float b) { return (a < b) ? b : a; }
KRT_DEV float selmin(float a, float b) { return (b < a) ? b : a; }
KRT_DEV bool isnan3(f3 v) { return isnan(v.x) || isnan(v.y) || isnan(v.z); }
KRT_DEV f3 ld3(const float4 &v) { return mk(v.x, v.y, v.z); }
KRT_DEV f3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

// glm::min / glm::max semantics
KRT_DEV float gmin(float x, float y) { return (y < x) ? y : x; }
KRT_DEV float gmax(float x, float y) { return (x < y) ? y : x; }

// ---------------------------------------------------------------------------
// RNG (random.cuh:14-58)
// ---------------------------------------------------------------------------
KRT_DEV void pcg3d(uint32_t &x, uint32_t &y, uint32_t &z) {
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    x += y * z;
    y += z * x;
    z += x * y;
    x ^= x >> 16u;
    y ^= y >> 16u;
    z ^= z >> 16u;
    x += y * z;
    y += z * x;
    z += x * y;
}

KRT_DEV f3 pcg3f(f3 &seed) {
    uint32_t x = __float_as_uint(seed.x), y = __float_as_uint(seed.y), z = __float_as_uint(seed.z);
    pcg3d(x, y, z);
    x = (x & 0x007fffffu) | 0x3f800000u;
    y = (y & 0x007fffffu) | 0x3f800000u;
    z = (z & 0x007fffffu) | 0x3f800000u;
    seed = mk(__uint_as_float(x) - 1.0f, __uint_as_float(y) - 1.0f, __uint_as_float(z) - 1.0f);
    return seed;
}

KRT_DEV float rand_uniform(f3 &seed) {
    seed = pcg3f(seed);
    return fract(seed.x);
}

KRT_DEV f3 rand_uniform_3f(f3 &seed) {
    seed = pcg3f(seed);
    return mk(fract(seed.x), fract(seed.y), fract(seed.z));
}

// Work item -> pixel of a grid-stride pixel loop, in 8x8 tiles taken in
// row-major tile order: the 64 lanes of a wavefront shade one tile, so their
// primary rays, G-buffer reads and light samples stay close on screen.  Items
// of a partial edge tile that fall outside the image return false.
struct TileOrder {
    uint32_t width, height, tiles_x, items;
    KRT_DEV TileOrder(uint32_t w, uint32_t h)
        : width(w), height(h), tiles_x((w + 7u) >> 3), items(((w + 7u) >> 3) * ((h + 7u) >> 3) * 64u) {}
    KRT_DEV bool pixel(uint32_t item, uint32_t &px, uint32_t &py) const {
        const uint32_t tile = item >> 6, t = item & 63u;
        px = (tile % tiles_x) * 8u + (t & 7u);
        py = (tile / tiles_x) * 8u + (t >> 3);
        return px < width && py < height;
    }
};

// Workgroups go to the chip's 8 XCDs round-robin (workgroup b to XCD b % 8),
// and each XCD has its own L2.  Per-pixel kernels take their tiles through
// this bijection of [0, gridDim.x), which hands XCD x one contiguous run of
// workgroups (a band of image rows), so that the neighbourhood reads of a
// tile (spatial reuse, reprojection) mostly hit lines its own L2 already
// holds instead of pulling them into all eight.
constexpr uint32_t kXcds = 8;
KRT_DEV uint32_t xcd_block() {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t q = G / kXcds, r = G % kXcds, x = b % kXcds, j = b / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// Seeds: three u32 words bit-cast into the float3 state.
KRT_DEV f3 pixel_seed(uint32_t px, uint32_t py, uint32_t sample, uint32_t user) {
    return mk(__uint_as_float(px), __uint_as_float(py), __uint_as_float(sample + 0x9E3779B9u * user));
}
KRT_DEV f3 reservoir_seed(uint32_t px, uint32_t py, uint32_t user) {
    f3 s = mk(__uint_as_float(px), __uint_as_float(py), __uint_as_float(0x5EED0000u ^ (0x9E3779B9u * user)));
    pcg3f(s);
    return s;
}

// cos(2 pi u), sin(2 pi u), u in [0,1): exact quarter-turn reduction + Taylor
// polynomials on [-pi/4, pi/4].
KRT_DEV void sincos_turn(float u, float &c, float &s) {
    float q = floorf(u * 4.0f + 0.5f);
    float r = u - q * 0.25f;
    float x = r * 6.28318530717958647692f;
    float x2 = x * x;
    float sp = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f +
                  x2 * (-0.000198412701f + x2 * 2.75573188e-06f))));
    float cp = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f +
                  x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    int qi = ((int)q) & 3;
    c = (qi == 0) ? cp : (qi == 1) ? -sp : (qi == 2) ? -cp : sp;
    s = (qi == 0) ? sp : (qi == 1) ? cp : (qi == 2) ? -sp : -cp;
}

KRT_DEV f3 rotate(f3 s, f3 n) {  // math.cuh:2653-2667
    f3 w = n;
    f3 a = mk(0.0f, 1.0f, 0.0f);
    if (fabsf(dot(w, a)) > 0.999f) a = mk(0.0f, 0.0f, 1.0f);
    f3 u = normalize(cross(w, a));
    f3 v = normalize(cross(w, u));
    return (u * s.x + v * s.y) + w * s.z;
}

// ---------------------------------------------------------------------------
// Materials and BSDF (brdf.cuh)
// ---------------------------------------------------------------------------
enum : int { eReflection = 1, eTransmission = 2, eDiffuse = 4 };
constexpr float PI_F = 3.14159265358979323846f;

struct Material {
    f3 diffuse, specular, emission;
    float roughness, refraction;
    int type;
};

// Device material table row: 4 x float4 = {diffuse, rough}, {specular, refr},
// {emission, type-as-int-bits}, {texture ids as uint bits: diffuse | normal
// << 16, specular | emission << 16, roughness | 0xFFFF << 16, any-texture
// flag}; a 16-bit id of kNoTex means no texture.
KRT_DEV Material load_material(const float4 *tab, uint32_t m) {
    float4 a = tab[4 * m], b = tab[4 * m + 1], c = tab[4 * m + 2];
    Material M;
    M.diffuse = mk(a.x, a.y, a.z);
    M.roughness = a.w;
    M.specular = mk(b.x, b.y, b.z);
    M.refraction = b.w;
    M.emission = mk(c.x, c.y, c.z);
    M.type = __float_as_int(c.w);
    return M;
}

constexpr uint32_t kNoTex = 0xFFFFu;
KRT_DEV uint4 material_tex(const float4 *tab, uint32_t m) {
    const float4 r = tab[4 * m + 3];
    return make_uint4(__float_as_uint(r.x), __float_as_uint(r.y), __float_as_uint(r.z), __float_as_uint(r.w));
}

// An RGBA8 texel (r in the low byte) read as byte / 255 per channel
// (cudaReadModeNormalizedFloat).
KRT_DEV float4 texel8(uint32_t t) {
    return make_float4((float)(t & 255u) / 255.0f, (float)((t >> 8) & 255u) / 255.0f,
                       (float)((t >> 16) & 255u) / 255.0f, (float)(t >> 24) / 255.0f);
}

struct SurfaceHit {
    Material mat;
    bool entering;
    f3 n, wo, x;
};

KRT_DEV float G1(f3 n, f3 v, const Material &mat) {  // brdf.cuh:180-192
    if (dot(v, n) <= 0.0f) return 0.0f;
    float alpha = mat.roughness;
    float c = clampf(dot(n, v), 0.0f, 0.999f);
    float c2 = c * c;
    float tan2 = (1.0f - c2) / c2;
    float denom = 1.0f + sqrtf(1.0f + alpha * alpha * tan2);
    return 2.0f / denom;
}

KRT_DEV f3 shlick_F(f3 wi, f3 h, const Material &mat) {  // brdf.cuh:203-207
    float x = 1.0f - dot(wi, h);
    float x2 = x * x;
    float k = x2 * x2 * x;
    return mat.specular + (1.0f - mat.specular) * k;
}

KRT_DEV float ggx_D(f3 n, f3 h, const Material &mat) {  // brdf.cuh:211-219
    float alpha = mat.roughness;
    float a2 = alpha * alpha;
    float c = clampf(dot(n, h), 0.0f, 1.0f);
    float k = c * c * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * k * k);
}

KRT_DEV f3 ggx_brdf(const Material &mat, f3 n, f3 wi, f3 wo) {  // brdf.cuh:337-353
    if (dot(wo, n) <= 0.0f) n = -n;
    f3 h = normalize(wi + wo);
    f3 f = shlick_F(wi, h, mat);
    float g = G1(n, wo, mat) * G1(n, wi, mat);
    float d = ggx_D(n, h, mat);
    f3 num = f * g * d;
    float denom = 4.0f * dot(wi, n) * dot(wo, n);
    return num / (denom + 1e-6f);
}

KRT_DEV float ggx_pdf(const Material &mat, f3 n, f3 wi, f3 wo) {  // brdf.cuh:357-382
    if (dot(wo, n) <= 0.0f) n = -n;
    if (dot(wi, n) <= 0.0f) return 0.0f;
    f3 h = normalize(wi + wo);
    float avg_Kd = (mat.diffuse.x + mat.diffuse.y + mat.diffuse.z) / 3.0f;
    float avg_Ks = (mat.specular.x + mat.specular.y + mat.specular.z) / 3.0f;
    float t = 1.0f;
    if (avg_Kd + avg_Ks > 0.0f) t = selmax(avg_Ks / (avg_Kd + avg_Ks), 0.25f);
    float term1 = dot(n, wi) / PI_F;
    float term2 = ggx_D(n, h, mat) * dot(n, h) / (4.0f * dot(wi, h));
    return (1.0f - t) * term1 + t * term2;
}

KRT_DEV f3 ggx_sample(const Material &mat, f3 n, f3 wo, f3 &seed) {  // brdf.cuh:386-433
    if (dot(wo, n) <= 0.0f) n = -n;
    float avg_Kd = (mat.diffuse.x + mat.diffuse.y + mat.diffuse.z) / 3.0f;
    float avg_Ks = (mat.specular.x + mat.specular.y + mat.specular.z) / 3.0f;
    float t = 1.0f;
    if (avg_Kd + avg_Ks > 0.0f) t = avg_Ks / (avg_Kd + avg_Ks);
    f3 eta = rand_uniform_3f(seed);
    float cphi, sphi;
    sincos_turn(eta.z, cphi, sphi);
    float ct, st;
    bool spec = eta.x < t;
    if (spec) {
        float k = sqrtf(eta.y / (1.0f - eta.y));
        float tn = k * mat.roughness;
        ct = 1.0f / sqrtf(1.0f + tn * tn);
        st = tn * ct;
    } else {
        ct = sqrtf(eta.y);
        st = sqrtf(1.0f - eta.y);
    }
    f3 v = rotate(mk(st * cphi, st * sphi, ct), n);
    return spec ? reflect(-wo, v) : v;
}

KRT_DEV f3 glass_sample(const SurfaceHit &sh, f3 wi) {  // brdf.cuh:144-175
    float eta = sh.mat.refraction;
    float cos_theta = selmin(dot(-wi, sh.n), 1.0f);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    bool reflect_ray = (eta * sin_theta > 1.0f);
    float ni_over_nt = 1.0f / eta;
    if (!sh.entering) ni_over_nt = eta;
    if (reflect_ray) return reflect(wi, sh.n);
    f3 v = wi, n = sh.n;
    float c = dot(-v, n);
    f3 par = ni_over_nt * (v + c * n);
    float len = length(par);
    float len2 = len * len;
    f3 perp = -sqrtf(fabsf(1.0f - len2)) * n;
    return par + perp;
}

KRT_DEV f3 brdf(const SurfaceHit &sh, f3 wi, int out) {  // brdf.cuh:437-449
    f3 diff = sh.mat.diffuse / PI_F;
    if (out & eTransmission) return diff + mk(0.0f, 0.0f, 0.0f);
    return diff + ggx_brdf(sh.mat, sh.n, wi, sh.wo);
}

KRT_DEV f3 eval_bsdf(const SurfaceHit &sh, f3 &wi, float &pdf, int &out, f3 &seed) {  // brdf.cuh:573-589
    if (sh.mat.type & eTransmission) {
        wi = glass_sample(sh, sh.wo);
        pdf = 1.0f;
        out = sh.mat.type;
        return sh.mat.specular;
    }
    wi = ggx_sample(sh.mat, sh.n, sh.wo, seed);
    out = eDiffuse;
    if (length(wi) < 1e-6f) {
        pdf = 0.0f;
        return mk(0.0f, 0.0f, 0.0f);
    }
    pdf = ggx_pdf(sh.mat, sh.n, wi, sh.wo);
    return brdf(sh, wi, out);
}

// ---------------------------------------------------------------------------
// Triangle lights (lighting.cuh:96-146).  Device row: 4 x float4 =
// {a, area}, {ab, 0}, {ac, 0}, {intensity, 0} plus a normal table; area and
// normal are computed once with the same expressions as the per-sample form.
// ---------------------------------------------------------------------------
struct TriLight {
    f3 a, ab, ac, intensity;
    float area;
    f3 normal;
};

KRT_DEV float light_area(f3 ab, f3 ac) { return length(cross(ab, ac)) * 0.5f; }
KRT_DEV f3 light_normal(f3 ab, f3 ac) { return normalize(cross(ab, ac)); }

KRT_DEV TriLight load_light(const float4 *tab, uint32_t i) {
    float4 a = tab[4 * i], b = tab[4 * i + 1], c = tab[4 * i + 2], d = tab[4 * i + 3];
    TriLight L;
    L.a = mk(a.x, a.y, a.z);
    L.area = a.w;
    L.ab = mk(b.x, b.y, b.z);
    L.normal.x = b.w;
    L.ac = mk(c.x, c.y, c.z);
    L.normal.y = c.w;
    L.intensity = mk(d.x, d.y, d.z);
    L.normal.z = d.w;
    return L;
}

KRT_DEV f3 sample_area_light(const TriLight &l, f3 &seed) {
    f3 r = pcg3f(seed);
    float u = fract(r.x), v = fract(r.y);
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    return (l.a + u * l.ab) + v * l.ac;
}

KRT_DEV unsigned pick_light(f3 &seed, unsigned total) {
    unsigned i = (unsigned)(rand_uniform(seed) * (float)total);
    return i < total ? i : total - 1;
}

// ---------------------------------------------------------------------------
// Environment map (ArmadaRTX::set_envmap, source/amadeus/armada.cu:86-96; the
// lookups of amadeus_path_tracer.cu:204-220, amadeus_restir.cu:566-577 and
// Ld_Environment / sample_direct, amadeus_common.cuh:215-252, 341-370).
// atan and asin are their Maclaurin series after argument reduction (the
// same expressions in the oracle, so both give the same bits):
//   atan t = t - t^3/3 + t^5/5 - ... for |t| <= tan(pi/8); above it
//   atan t = pi/4 + atan((t - 1)/(t + 1)); above 1, atan t = pi/2 - atan(1/t);
//   asin x = sum (2k)!/(4^k (k!)^2 (2k+1)) x^(2k+1) for |x| <= 1/2; above it
//   asin x = pi/2 - 2 asin(sqrt((1 - x)/2)).
// ---------------------------------------------------------------------------
constexpr float kPiF = 0x1.921fb6p+1f;
constexpr float kHalfPiF = 0x1.921fb6p+0f;
constexpr float kQuarterPiF = 0x1.921fb6p-1f;
constexpr float kTwoPiF = 0x1.921fb6p+2f;
constexpr float kFourPiF = 0x1.921fb6p+3f;
constexpr float kTanPi8F = 0x1.a8279ap-2f;
constexpr float kEnvDistance = 10000.0f;  // WORLD_RADIUS, amadeus_common.cuh:219

KRT_DEV float atan_series(float t) {  // |t| <= tan(pi/8)
    const float t2 = t * t;
    float p = 0x1.47ae14p-5f;  // 1/25
    p = p * t2 + -0x1.642c86p-5f;
    p = p * t2 + 0x1.861862p-5f;
    p = p * t2 + -0x1.af286cp-5f;
    p = p * t2 + 0x1.e1e1e2p-5f;
    p = p * t2 + -0x1.111112p-4f;
    p = p * t2 + 0x1.3b13b2p-4f;
    p = p * t2 + -0x1.745d18p-4f;
    p = p * t2 + 0x1.c71c72p-4f;
    p = p * t2 + -0x1.24924ap-3f;
    p = p * t2 + 0x1.99999ap-3f;
    p = p * t2 + -0x1.555556p-2f;
    return t + (t * t2) * p;
}

KRT_DEV float atan_unit(float t) {  // 0 <= t <= 1
    return t > kTanPi8F ? kQuarterPiF + atan_series((t - 1.0f) / (t + 1.0f)) : atan_series(t);
}

KRT_DEV float krt_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) r = signbit(x) ? kPiF : 0.0f;
    else {
        r = ay <= ax ? atan_unit(ay / ax) : kHalfPiF - atan_unit(ax / ay);
        if (signbit(x)) r = kPiF - r;
    }
    return signbit(y) ? -r : r;
}

KRT_DEV float asin_series(float x) {  // |x| <= 1/2
    const float x2 = x * x;
    float p = 0x1.df3bd4p-8f;  // 88179/12058624
    p = p * x2 + 0x1.12ef3cp-7f;
    p = p * x2 + 0x1.3fde50p-7f;
    p = p * x2 + 0x1.7a8788p-7f;
    p = p * x2 + 0x1.c9999ap-7f;
    p = p * x2 + 0x1.1c4ec4p-6f;
    p = p * x2 + 0x1.6e8ba2p-6f;
    p = p * x2 + 0x1.f1c71cp-6f;
    p = p * x2 + 0x1.6db6dcp-5f;
    p = p * x2 + 0x1.333334p-4f;
    p = p * x2 + 0x1.555556p-3f;
    return x + (x * x2) * p;
}

KRT_DEV float krt_asin(float x) {  // x clamped to [-1, 1]
    const float a = fminf(fabsf(x), 1.0f);
    const float r = a <= 0.5f ? asin_series(a) : kHalfPiF - 2.0f * asin_series(sqrtf((1.0f - a) * 0.5f));
    return signbit(x) ? -r : r;
}

struct EnvMap {  // RGBA float texels, row-major, width x height; tex == nullptr: none
    const float4 *tex;
    uint32_t w, h;
};

KRT_DEV int env_wrap(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// tex2D <float4> with normalised coordinates, bilinear filtering and wrap
// addressing (the texture descriptors of include/cuda/interop.cuh:96-135),
// filtered in FP32 (the texture unit's 8-bit weights are not restated).
KRT_DEV f3 env_fetch(const EnvMap &env, float u, float v) {
    if (!(u >= -1.0f && u <= 2.0f && v >= -1.0f && v <= 2.0f)) return mk(0.0f, 0.0f, 0.0f);  // NaN direction
    const float x = u * (float)env.w - 0.5f, y = v * (float)env.h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int ix = (int)fx, iy = (int)fy, W = (int)env.w, H = (int)env.h;
    const int i0 = env_wrap(ix, W), i1 = env_wrap(ix + 1, W), j0 = env_wrap(iy, H), j1 = env_wrap(iy + 1, H);
    const f3 t00 = ld3(env.tex[(size_t)j0 * env.w + i0]), t10 = ld3(env.tex[(size_t)j0 * env.w + i1]);
    const f3 t01 = ld3(env.tex[(size_t)j1 * env.w + i0]), t11 = ld3(env.tex[(size_t)j1 * env.w + i1]);
    const float a0 = 1.0f - a, b0 = 1.0f - b;
    return (t00 * (a0 * b0) + t10 * (a * b0)) + (t01 * (a0 * b) + t11 * (a * b));
}

// Texture coordinates of a direction: u = atan2(x, z) / 2 pi + 1/2,
// v = asin(y) / pi + 1/2.
KRT_DEV void env_uv(f3 d, float &u, float &v) {
    u = krt_atan2(d.x, d.z) / kTwoPiF + 0.5f;
    v = krt_asin(d.y) / kPiF + 0.5f;
}

// Radiance of the environment seen along d: the path tracer's miss program
// flips v (amadeus_path_tracer.cu:215-216), ReSTIR's and the light sampling do
// not (amadeus_restir.cu:573-574, amadeus_common.cuh:237-240).
KRT_DEV f3 env_radiance(const EnvMap &env, f3 d, bool flip_v) {
    float u, v;
    env_uv(d, u, v);
    return env_fetch(env, u, flip_v ? 1.0f - v : v);
}

// The environment light's direction sample (Ld_Environment /
// sample_direct): one pcg3f step, theta = acos(sqrt(1 - u)) about the world
// +z axis, phi = 2 pi u'; with cos(theta) = sqrt(1 - u) and sin(theta) =
// sqrt(1 - cos^2(theta)).  Its pdf is taken as 1 / (4 pi) by the reference.
KRT_DEV f3 env_sample_dir(f3 &seed) {
    seed = rand_uniform_3f(seed);
    const float ct = sqrtf(1.0f - fract(seed.x));
    const float st = sqrtf(1.0f - ct * ct);
    float cp, sp;
    sincos_turn(fract(seed.y), cp, sp);
    return mk(st * cp, st * sp, ct);
}
constexpr float kEnvPdf = 0x1.45f306p-4f;  // 1 / (4 pi)

// Accumulation on a glm::vec4 (amadeus_path_tracer.cu:59-70): true division.
KRT_DEV float4 accumulate4(float4 dst, float4 s, bool acc, float count) {
    if (!acc) return s;
    float c1 = count + 1.0f;
    return make_float4((dst.x * count + s.x) / c1, (dst.y * count + s.y) / c1, (dst.z * count + s.z) / c1,
                       (dst.w * count + s.w) / c1);
}

// (int)((double)x + 0.5) of the reference's reprojection (amadeus_restir.cu:
// 202-203, a double literal) for 0 <= x < 2^23 without FP64: the double sum
// is exact, so it equals floor(x) + (frac(x) >= 0.5), and x - floor(x) is
// exact in float for such x.
KRT_DEV int round_half_up(float x) {
    const int i = (int)x;
    return i + ((x - (float)i) >= 0.5f ? 1 : 0);
}

// Wave-level counter add (one atomic per wavefront).
KRT_DEV void wave_count(unsigned long long *ctr, unsigned v) {
    if (!ctr) return;
    unsigned long long s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(ctr, s);
}

}  // namespace krt

// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Scalar restatement of the reference renderer's device math (vector helpers,
// pcg RNG, GGX/Glass BSDF, triangle lights).  Only tests/, the smoke check and
// bench.py's cpu_baseline leg may load anything built from oracle/.  Nothing in
// the product path (artemis_amd/, include/) may include or link it.
//
// Build: g++ -O2 -ffp-contract=off (no fast-math), so every a*b+c below is two
// IEEE roundings, exactly as the HIP kernels compute it.
//
// Parity pins for each function live in tests/ (see the header of oracle.h).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct f3 {
    float x, y, z;
};

static inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
// include/cuda/math.cuh:401-470 (component-wise operators; a/s = a * (1/s))
static inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 operator*(float s, f3 a) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 operator/(f3 a, float s) {
    float inv = 1.0f / s;
    return a * inv;
}
static inline f3 operator+(float a, f3 b) { return mk(a + b.x, a + b.y, a + b.z); }
static inline f3 operator-(float a, f3 b) { return mk(a - b.x, a - b.y, a - b.z); }
// include/cuda/math.cuh:508-530
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length(f3 v) { return sqrtf(dot(v, v)); }
static inline f3 normalize(f3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
// include/cuda/math.cuh:539
static inline f3 reflect(f3 i, f3 n) { return i - (2.0f * n) * dot(n, i); }
// include/cuda/math.cuh:2604-2607
static inline float fract(float x) { return x - floorf(x); }
// Clamp and min/max as compare-selects (identical on CPU and GPU, NaN passes).
static inline float clampf(float x, float a, float b) {
    x = (x < a) ? a : x;
    return (x > b) ? b : x;
}
static inline float selmax(float a, float b) { return (a < b) ? b : a; }
static inline float selmin(float a, float b) { return (b < a) ? b : a; }
static inline bool isnan3(f3 v) { return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z); }

// glm::min / glm::max scalar semantics: (y < x) ? y : x and (x < y) ? y : x.
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
static inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// RNG — include/cuda/random.cuh:14-58 (pcg3d of Jarzynski & Olano, JCGT 2020).
// The float3 seed's BITS are hashed; the state becomes the [0,1) floats.
// ---------------------------------------------------------------------------
static inline void pcg3d(uint32_t v[3]) {
    for (int i = 0; i < 3; i++) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    for (int i = 0; i < 3; i++) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
}

static inline f3 pcg3f(f3 &seed) {
    uint32_t v[3] = {fbits(seed.x), fbits(seed.y), fbits(seed.z)};
    pcg3d(v);
    for (int i = 0; i < 3; i++) v[i] = (v[i] & 0x007fffffu) | 0x3f800000u;
    seed = mk(bitsf(v[0]) - 1.0f, bitsf(v[1]) - 1.0f, bitsf(v[2]) - 1.0f);
    return seed;
}

static inline float rand_uniform(f3 &seed) {
    seed = pcg3f(seed);
    return fract(seed.x);
}

static inline f3 rand_uniform_3f(f3 &seed) {
    seed = pcg3f(seed);
    return mk(fract(seed.x), fract(seed.y), fract(seed.z));
}

// Pinned seed scheme (replaces the wall-clock / sinf seed of
// source/optix/amadeus_path_tracer.cu:99-107): three u32 words bit-cast into
// the float3 seed.
static inline f3 seed_from_words(uint32_t a, uint32_t b, uint32_t c) {
    return mk(bitsf(a), bitsf(b), bitsf(c));
}
static inline f3 pixel_seed(uint32_t px, uint32_t py, uint32_t sample, uint32_t user) {
    return seed_from_words(px, py, sample + 0x9E3779B9u * user);
}
// Per-pixel reservoir seed at load time (replaces the std::random_device seeds
// of include/amadeus/reservoir.cuh:29-36): one pcg3f step of a fixed word.
static inline f3 reservoir_seed(uint32_t px, uint32_t py, uint32_t user) {
    f3 s = seed_from_words(px, py, 0x5EED0000u ^ (0x9E3779B9u * user));
    pcg3f(s);
    return s;
}

// ---------------------------------------------------------------------------
// cos(2*pi*u), sin(2*pi*u) for u in [0,1): exact quarter-turn reduction, then
// Taylor polynomials on [-pi/4, pi/4] (truncation < 2e-9).  Used wherever the
// reference evaluates sin/cos of 2*pi*rand (brdf.cuh:405-428,
// amadeus_restir.cu:304-308); identical op order on CPU and GPU.
// ---------------------------------------------------------------------------
static inline void sincos_turn(float u, float &c, float &s) {
    float q = floorf(u * 4.0f + 0.5f);
    float r = u - q * 0.25f;  // exact (Sterbenz)
    float x = r * 6.28318530717958647692f;
    float x2 = x * x;
    float sp = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f +
                  x2 * (-0.000198412701f + x2 * 2.75573188e-06f))));
    float cp = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f +
                  x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    int qi = ((int)q) & 3;
    if (qi == 0) { c = cp; s = sp; }
    else if (qi == 1) { c = -sp; s = cp; }
    else if (qi == 2) { c = -cp; s = -sp; }
    else { c = sp; s = -cp; }
}

// include/cuda/math.cuh:2653-2667
static inline f3 rotate(f3 s, f3 n) {
    f3 w = n;
    f3 a = mk(0.0f, 1.0f, 0.0f);
    if (fabsf(dot(w, a)) > 0.999f) a = mk(0.0f, 0.0f, 1.0f);
    f3 u = normalize(cross(w, a));
    f3 v = normalize(cross(w, u));
    return (u * s.x + v * s.y) + w * s.z;
}

// ---------------------------------------------------------------------------
// Materials and BSDF — include/cuda/brdf.cuh
// ---------------------------------------------------------------------------
enum : int {
    eNone = 0,
    eReflection = 1 << 0,
    eTransmission = 1 << 1,
    eDiffuse = 1 << 2,
};

static const float PI_F = 3.14159265358979323846f;

struct Material {  // include/cuda/material.cuh:13-23 (the fields the path reads)
    f3 diffuse, specular, emission;
    float roughness, refraction;
    int type;
};

struct SurfaceHit {  // brdf.cuh:67-74
    Material mat;
    bool entering;
    f3 n, wo, x;
};

// brdf.cuh:180-192.  tan^2(acos c) is evaluated as (1 - c^2) / c^2, the exact
// identity, instead of the CUDA fast-intrinsic __tanf of acosf.
static inline float G1(f3 n, f3 v, const Material &mat) {
    if (dot(v, n) <= 0.0f) return 0.0f;
    float alpha = mat.roughness;
    float c = clampf(dot(n, v), 0.0f, 0.999f);
    float c2 = c * c;
    float tan2 = (1.0f - c2) / c2;
    float denom = 1.0f + sqrtf(1.0f + alpha * alpha * tan2);
    return 2.0f / denom;
}

static inline float G(f3 n, f3 wi, f3 wo, const Material &mat) { return G1(n, wo, mat) * G1(n, wi, mat); }

// brdf.cuh:203-207; powf(x, 5) as x^2 * x^2 * x.
static inline f3 shlick_F(f3 wi, f3 h, const Material &mat) {
    float x = 1.0f - dot(wi, h);
    float x2 = x * x;
    float k = x2 * x2 * x;
    return mat.specular + (1.0f - mat.specular) * k;
}

// brdf.cuh:211-219.  alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2) rewritten with
// the identity cos^4 (alpha^2 + tan^2)^2 = (c^2 (alpha^2 - 1) + 1)^2.
static inline float ggx_D(f3 n, f3 h, const Material &mat) {
    float alpha = mat.roughness;
    float a2 = alpha * alpha;
    float c = clampf(dot(n, h), 0.0f, 1.0f);
    float k = c * c * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * k * k);
}

// brdf.cuh:337-353
static inline f3 ggx_brdf(const Material &mat, f3 n, f3 wi, f3 wo) {
    if (dot(wo, n) <= 0.0f) n = -n;
    f3 h = normalize(wi + wo);
    f3 f = shlick_F(wi, h, mat);
    float g = G(n, wi, wo, mat);
    float d = ggx_D(n, h, mat);
    f3 num = f * g * d;
    float denom = 4.0f * dot(wi, n) * dot(wo, n);
    return num / (denom + 1e-6f);
}

// brdf.cuh:357-382 (pdf mixture with t = max(Ks/(Kd+Ks), 0.25): quirk kept)
static inline float ggx_pdf(const Material &mat, f3 n, f3 wi, f3 wo) {
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

// brdf.cuh:386-433 (lobe choice t = Ks/(Kd+Ks), no 0.25 floor: quirk kept).
// theta = atan(k alpha) enters only through cos/sin: cos = 1/sqrt(1+tan^2),
// sin = tan*cos.  Diffuse theta = acos(sqrt(u)): cos = sqrt(u), sin = sqrt(1-u).
static inline f3 ggx_sample(const Material &mat, f3 n, f3 wo, f3 &seed) {
    if (dot(wo, n) <= 0.0f) n = -n;
    float avg_Kd = (mat.diffuse.x + mat.diffuse.y + mat.diffuse.z) / 3.0f;
    float avg_Ks = (mat.specular.x + mat.specular.y + mat.specular.z) / 3.0f;
    float t = 1.0f;
    if (avg_Kd + avg_Ks > 0.0f) t = avg_Ks / (avg_Kd + avg_Ks);
    f3 eta = rand_uniform_3f(seed);
    float cphi, sphi;
    sincos_turn(eta.z, cphi, sphi);
    if (eta.x < t) {
        float k = sqrtf(eta.y / (1.0f - eta.y));
        float tn = k * mat.roughness;
        float ct = 1.0f / sqrtf(1.0f + tn * tn);
        float st = tn * ct;
        f3 h = mk(st * cphi, st * sphi, ct);
        h = rotate(h, n);
        return reflect(-wo, h);
    }
    float ct = sqrtf(eta.y);
    float st = sqrtf(1.0f - eta.y);
    f3 s = mk(st * cphi, st * sphi, ct);
    return rotate(s, n);
}

// brdf.cuh:144-175 (Glass::sample; brdf/pdf of Glass are 0)
static inline f3 glass_refract(f3 v, f3 n, float ni_over_nt) {
    float cos_theta = dot(-v, n);
    f3 r_out_parallel = ni_over_nt * (v + cos_theta * n);
    float r_out_length = length(r_out_parallel);
    float r_out_length_squared = r_out_length * r_out_length;
    f3 r_out_perp = -sqrtf(fabsf(1.0f - r_out_length_squared)) * n;
    return r_out_parallel + r_out_perp;
}

static inline f3 glass_sample(const SurfaceHit &sh, f3 wi) {
    float eta = sh.mat.refraction;
    float cos_theta = selmin(dot(-wi, sh.n), 1.0f);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    bool reflect_ray = (eta * sin_theta > 1.0f);
    float ni_over_nt = 1.0f / eta;
    if (!sh.entering) ni_over_nt = eta;
    return reflect_ray ? reflect(wi, sh.n) : glass_refract(wi, sh.n, ni_over_nt);
}

// brdf.cuh:437-449
static inline f3 brdf(const SurfaceHit &sh, f3 wi, int out) {
    f3 diff = sh.mat.diffuse / PI_F;
    if (out & eTransmission) return diff + mk(0.0f, 0.0f, 0.0f);
    return diff + ggx_brdf(sh.mat, sh.n, wi, sh.wo);
}

// brdf.cuh:573-589 with eval<GGX> :477-492.  A degenerate sample (|wi| < 1e-6)
// returns f = 0 and pdf = 0 (the reference leaves pdf unset there).
static inline f3 eval_bsdf(const SurfaceHit &sh, f3 &wi, float &pdf, int &out, f3 &seed) {
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
// Triangle lights — include/optix/lighting.cuh:96-146
// ---------------------------------------------------------------------------
struct TriLight {
    f3 a, ab, ac, intensity;
};

static inline float light_area(const TriLight &l) { return length(cross(l.ab, l.ac)) * 0.5f; }
static inline f3 light_normal(const TriLight &l) { return normalize(cross(l.ab, l.ac)); }

static inline f3 sample_area_light(const TriLight &l, f3 &seed) {
    f3 r = pcg3f(seed);
    float u = fract(r.x), v = fract(r.y);
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    return (l.a + u * l.ab) + v * l.ac;
}

// ---------------------------------------------------------------------------
// Environment map (ArmadaRTX::set_envmap, source/amadeus/armada.cu:86-96; the
// miss programs amadeus_path_tracer.cu:204-220 and amadeus_restir.cu:566-577;
// Ld_Environment and sample_direct, amadeus_common.cuh:215-252, 341-370).
// CUDA's atan2f / asinf are not restated: atan and asin here are their
// Maclaurin series after argument reduction,
//   atan t = t - t^3/3 + ... - t^23/23 + t^25/25 for |t| <= tan(pi/8);
//   atan t = pi/4 + atan((t - 1)/(t + 1)) above it; pi/2 - atan(1/t) above 1;
//   asin x = sum_k (2k)!/(4^k (k!)^2 (2k+1)) x^(2k+1), k <= 11, for |x| <= 1/2;
//   asin x = pi/2 - 2 asin(sqrt((1 - x)/2)) above it.
// The texture fetch is tex2D with normalised coordinates, bilinear filtering
// and wrap addressing (include/cuda/interop.cuh:96-135), in FP32.
// ---------------------------------------------------------------------------
static const float kPiF = 0x1.921fb6p+1f;
static const float kHalfPiF = 0x1.921fb6p+0f;
static const float kQuarterPiF = 0x1.921fb6p-1f;
static const float kTwoPiF = 0x1.921fb6p+2f;
static const float kFourPiF = 0x1.921fb6p+3f;
static const float kTanPi8F = 0x1.a8279ap-2f;
static const float kEnvDistance = 10000.0f;        // WORLD_RADIUS, amadeus_common.cuh:219
static const float kEnvPdf = 0x1.45f306p-4f;       // 1 / (4 pi), amadeus_common.cuh:245

static inline float atan_series(float t) {
    static const float c[12] = {0x1.47ae14p-5f, -0x1.642c86p-5f, 0x1.861862p-5f, -0x1.af286cp-5f,
                                0x1.e1e1e2p-5f, -0x1.111112p-4f, 0x1.3b13b2p-4f, -0x1.745d18p-4f,
                                0x1.c71c72p-4f, -0x1.24924ap-3f, 0x1.99999ap-3f, -0x1.555556p-2f};
    const float t2 = t * t;
    float p = c[0];
    for (int k = 1; k < 12; k++) p = p * t2 + c[k];
    return t + (t * t2) * p;
}

static inline float atan_unit(float t) {
    return t > kTanPi8F ? kQuarterPiF + atan_series((t - 1.0f) / (t + 1.0f)) : atan_series(t);
}

static inline float orc_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = std::signbit(x) ? kPiF : 0.0f;
    } else {
        r = ay <= ax ? atan_unit(ay / ax) : kHalfPiF - atan_unit(ax / ay);
        if (std::signbit(x)) r = kPiF - r;
    }
    return std::signbit(y) ? -r : r;
}

static inline float asin_series(float x) {
    static const float c[11] = {0x1.df3bd4p-8f, 0x1.12ef3cp-7f, 0x1.3fde50p-7f, 0x1.7a8788p-7f,
                                0x1.c9999ap-7f, 0x1.1c4ec4p-6f, 0x1.6e8ba2p-6f, 0x1.f1c71cp-6f,
                                0x1.6db6dcp-5f, 0x1.333334p-4f, 0x1.555556p-3f};
    const float x2 = x * x;
    float p = c[0];
    for (int k = 1; k < 11; k++) p = p * x2 + c[k];
    return x + (x * x2) * p;
}

static inline float orc_asin(float x) {
    const float a = fminf(fabsf(x), 1.0f);
    const float r = a <= 0.5f ? asin_series(a) : kHalfPiF - 2.0f * asin_series(sqrtf((1.0f - a) * 0.5f));
    return std::signbit(x) ? -r : r;
}

struct EnvMap {  // RGBA float texels, row-major, w x h; tex == nullptr: none
    const float *tex = nullptr;
    uint32_t w = 0, h = 0;
};

static inline int env_wrap(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

static inline f3 env_texel(const EnvMap &env, int i, int j) {
    const float *t = env.tex + 4 * ((size_t)j * env.w + (size_t)i);
    return mk(t[0], t[1], t[2]);
}

static inline f3 env_fetch(const EnvMap &env, float u, float v) {
    if (!(u >= -1.0f && u <= 2.0f && v >= -1.0f && v <= 2.0f)) return mk(0, 0, 0);  // NaN direction
    const float x = u * (float)env.w - 0.5f, y = v * (float)env.h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int ix = (int)fx, iy = (int)fy, W = (int)env.w, H = (int)env.h;
    const int i0 = env_wrap(ix, W), i1 = env_wrap(ix + 1, W), j0 = env_wrap(iy, H), j1 = env_wrap(iy + 1, H);
    const f3 t00 = env_texel(env, i0, j0), t10 = env_texel(env, i1, j0);
    const f3 t01 = env_texel(env, i0, j1), t11 = env_texel(env, i1, j1);
    const float a0 = 1.0f - a, b0 = 1.0f - b;
    return (t00 * (a0 * b0) + t10 * (a * b0)) + (t01 * (a0 * b) + t11 * (a * b));
}

// u = atan2(x, z) / 2 pi + 1/2, v = asin(y) / pi + 1/2
static inline void env_uv(f3 d, float &u, float &v) {
    u = orc_atan2(d.x, d.z) / kTwoPiF + 0.5f;
    v = orc_asin(d.y) / kPiF + 0.5f;
}

// the path tracer's miss program reads (u, 1 - v); ReSTIR's and the light
// sampling read (u, v)
static inline f3 env_radiance(const EnvMap &env, f3 d, bool flip_v) {
    float u, v;
    env_uv(d, u, v);
    return env_fetch(env, u, flip_v ? 1.0f - v : v);
}

// Material textures (calculate_material / calculate_normal,
// amadeus_common.cuh:74-187): tex2D <float4> of an RGBA8 image
// (import_vulkan_texture, include/cuda/interop.cuh:11-51: read as byte / 255,
// normalised coordinates, bilinear, wrap addressing), filtered in FP32;
// coordinates beyond +-2^16 (or NaN) read zero.
struct Texture {
    const uint32_t *texels = nullptr;  // r in the low byte, row-major
    uint32_t w = 0, h = 0;
};

static inline void texel_rgba(uint32_t t, float out[4]) {
    for (int k = 0; k < 4; k++) out[k] = (float)((t >> (8 * k)) & 255u) / 255.0f;
}

static inline void tex_fetch(const Texture &tx, float u, float v, float out[4]) {
    if (!(std::fabs(u) <= 65536.0f && std::fabs(v) <= 65536.0f)) {
        for (int k = 0; k < 4; k++) out[k] = 0.0f;
        return;
    }
    const float x = u * (float)tx.w - 0.5f, y = v * (float)tx.h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int ix = (int)fx, iy = (int)fy, W = (int)tx.w, H = (int)tx.h;
    const int i0 = env_wrap(ix, W), i1 = env_wrap(ix + 1, W), j0 = env_wrap(iy, H), j1 = env_wrap(iy + 1, H);
    float t00[4], t10[4], t01[4], t11[4];
    texel_rgba(tx.texels[(size_t)j0 * tx.w + i0], t00);
    texel_rgba(tx.texels[(size_t)j0 * tx.w + i1], t10);
    texel_rgba(tx.texels[(size_t)j1 * tx.w + i0], t01);
    texel_rgba(tx.texels[(size_t)j1 * tx.w + i1], t11);
    const float a0 = 1.0f - a, b0 = 1.0f - b;
    const float w00 = a0 * b0, w10 = a * b0, w01 = a0 * b, w11 = a * b;
    for (int k = 0; k < 4; k++) out[k] = (t00[k] * w00 + t10[k] * w10) + (t01[k] * w01 + t11[k] * w11);
}

// Ld_Environment's direction: one pcg3f step; theta = acos(sqrt(1 - u)) about
// world +z (cos = sqrt(1 - u), sin = sqrt(1 - cos^2)), phi = 2 pi u'
static inline f3 env_sample_dir(f3 &seed) {
    seed = rand_uniform_3f(seed);
    const float ct = sqrtf(1.0f - fract(seed.x));
    const float st = sqrtf(1.0f - ct * ct);
    float cp, sp;
    sincos_turn(fract(seed.y), cp, sp);
    return mk(st * cp, st * sp, ct);
}

// Möller–Trumbore: shaders/rt/modules/intersect.glsl:14-33 ==
// include/optix/lighting.cuh:34-55.  Returns t, or -1 on a miss.
static inline float intersect_tri(f3 v1, f3 v2, f3 v3, f3 o, f3 d, float &b1, float &b2) {
    f3 e1 = v2 - v1;
    f3 e2 = v3 - v1;
    f3 s1 = cross(d, e2);
    float divisor = dot(s1, e1);
    if (divisor == 0.0f) return -1.0f;
    f3 s = o - v1;
    float inv = 1.0f / divisor;
    b1 = dot(s, s1) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return -1.0f;
    f3 s2 = cross(s, e1);
    b2 = dot(d, s2) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return -1.0f;
    return dot(e2, s2) * inv;
}

}  // namespace orc

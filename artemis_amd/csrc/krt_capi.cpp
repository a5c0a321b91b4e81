// C ABI of include/krt.h and the ArmadaRTX / AttachmentRTX state behind it.
//
// ArmadaRTX::render (source/amadeus/armada.cu:733-780): switch attachment
// (unload previous, load new), reset the sample count when not accumulating or
// when the acceleration structure changed, render, samples++.
// Attachment options follow AttachmentRTX::set_option (armada.cuh:122-126):
// unknown names are ignored, get_option of an unknown name is empty.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "krt_internal.h"
#include "krt_serve.h"

using namespace krt;

namespace krt {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

SceneView make_view(const Scene &s) {
    return SceneView{s.pos.as<float>(),   s.nrm.as<float>(),    s.idx.as<uint32_t>(),
                     s.tri_mat.as<uint32_t>(), s.mats.as<float4>(), s.lights.as<float4>(),
                     s.nl,                s.accel.tri.as<float4>(), s.accel.nodes.as<float4>(),
                     s.accel.wide.as<float4>(), s.tlas.wide.as<float4>(), s.inst_rec.as<float4>(),
                     s.textured ? s.uv.as<float2>() : nullptr,
                     s.textured && s.tan.p ? s.tan.as<float>() : nullptr,
                     TexSet{s.texels.as<uint32_t>(), s.tex_desc.as<uint4>()}};
}

}  // namespace krt

// ---------------------------------------------------------------------------
// Attachment and armada state
// ---------------------------------------------------------------------------
// A radiance-cache network: owned (krt_mlp_create) or a borrowed view of an
// attachment's cache (krt_attachment_cache; krt_mlp_destroy ignores it).
struct krt_mlp {
    Context *ctx = nullptr;
    Mlp *net = nullptr;
    std::unique_ptr<Mlp> owned;
};

struct krt_attachment {
    Context *ctx = nullptr;
    std::string kind;
    bool loaded = false;
    uint32_t user_seed = 0;
    int jitter = 1;  // path tracer pixel jitter (option "pixel_jitter")
    RestirOptions ropt;
    RestirState rst;
    GiState gst;
    MambaOptions mopt;
    MambaState mst;
    CacheOptions copt;
    CacheState cst;
    std::unique_ptr<krt_mlp> cache_handle;  // borrowed view of cst.net for krt_attachment_cache
    uint32_t w = 0, h = 0;

    void load(uint32_t width, uint32_t height) {  // restir.cuh:203-211
        w = width;
        h = height;
        if (kind == "restir") restir_alloc(*ctx, rst, width, height, user_seed);
        if (kind == "restir_gi") gi_alloc(*ctx, gst, width, height, user_seed);
        if (kind == "mamba") mamba_alloc(*ctx, mst, width, height, user_seed);
        loaded = true;
    }
    void unload() {  // restir.cuh:213-220
        rst = RestirState();
        gst = GiState();
        mst = MambaState();
        cache_handle.reset();
        cst = CacheState();
        loaded = false;
    }
    bool set_option(const std::string &f, const krt_option &v) {
        auto as_int = [&](int &dst) {
            if (v.kind == KRT_OPT_INT) dst = v.i;
            else if (v.kind == KRT_OPT_FLOAT) dst = (int)v.f;
            else if (v.kind == KRT_OPT_BOOL) dst = v.b ? 1 : 0;
            else return false;
            return true;
        };
        int tmp;
        if (f == "seed") {
            if (!as_int(tmp)) return false;
            user_seed = (uint32_t)tmp;
            return true;
        }
        if (kind == "path_tracer" && f == "pixel_jitter") return as_int(jitter);
        if (kind == "mamba") {  // mamba.cu temporal_reuse / spatial_reuse
            if (f == "temporal") return as_int(mopt.temporal);
            if (f == "spatial") return as_int(mopt.spatial);
            return true;
        }
        if (kind == "path_tracer" || kind == "restir") {
            if (f == "radiance_cache") return as_int(copt.enabled);
            if (f == "cache_depth") return as_int(copt.depth);
            if (f == "train_samples") return as_int(copt.train_samples);
            if (f == "cache_training") return as_int(copt.training);
            if (f == "cache_emission") return as_int(copt.learn_emission);
        }
        if (kind != "restir" && kind != "restir_gi") return true;  // unknown option names are ignored
        if (f == "candidates") return as_int(ropt.candidates);
        if (f == "m_cap") return as_int(ropt.m_cap);
        if (f == "spatial_passes") {
            // ReSTIR DI shades and accumulates the colour in its last spatial
            // pass (amadeus_restir.cu:420-433; the reference always runs two),
            // so it needs at least one; ReSTIR GI also shades in its temporal pass.
            if (!as_int(tmp) || tmp < (kind == "restir" ? 1 : 0)) return false;
            ropt.spatial_passes = tmp;
            return true;
        }
        if (f == "neighbours") return as_int(ropt.neighbours);
        if (f == "radius") return as_int(ropt.radius);
        return true;
    }
    krt_option get_option(const std::string &f) const {
        krt_option o{};
        o.kind = KRT_OPT_INT;
        if (f == "seed") o.i = (int)user_seed;
        else if (kind == "path_tracer" && f == "pixel_jitter") o.i = jitter;
        else if (kind == "mamba" && f == "temporal") o.kind = KRT_OPT_BOOL, o.b = mopt.temporal;
        else if (kind == "mamba" && f == "spatial") o.kind = KRT_OPT_BOOL, o.b = mopt.spatial;
        else if (kind == "mamba") o.kind = KRT_OPT_NONE;
        else if ((kind == "path_tracer" || kind == "restir") && f == "radiance_cache") o.i = copt.enabled;
        else if ((kind == "path_tracer" || kind == "restir") && f == "cache_depth") o.i = copt.depth;
        else if ((kind == "path_tracer" || kind == "restir") && f == "train_samples") o.i = copt.train_samples;
        else if ((kind == "path_tracer" || kind == "restir") && f == "cache_training") o.i = copt.training;
        else if ((kind == "path_tracer" || kind == "restir") && f == "cache_emission") o.i = copt.learn_emission;
        else if (kind != "path_tracer" && f == "candidates") o.i = ropt.candidates;
        else if (kind != "path_tracer" && f == "m_cap") o.i = ropt.m_cap;
        else if (kind != "path_tracer" && f == "spatial_passes") o.i = ropt.spatial_passes;
        else if (kind != "path_tracer" && f == "neighbours") o.i = ropt.neighbours;
        else if (kind != "path_tracer" && f == "radius") o.i = ropt.radius;
        else o.kind = KRT_OPT_NONE;
        return o;
    }
};

struct krt_armada {
    Context *ctx = nullptr;
    uint32_t w = 0, h = 0;
    DevBuf color, normal, albedo, position;
    DevBuf ldr;  // RGBA8 staging of krt_armada_read_ldr
    std::map<std::string, krt_attachment *> attachments;
    std::string active, previous;
    int max_depth = 2;  // armada.cu:56
    int samples = 0;
    krt_handle last_scene = 0;
    uint64_t last_version = 0;
    DevBuf ctr;  // unsigned long long x 16: closest, shadow x3 passes, nodes x3 passes, tris x3 passes
    krt_frame_stats stats{};
    DevBuf acc64;  // FP64 sample sums (path tracer), enabled on request
    bool f64 = false;
    uint64_t f64_frames = 0;
    DevBuf env;  // environment map texels (float4, row-major)
    uint32_t env_w = 0, env_h = 0;
    bool env_on = false;
};

struct krt_context {
    Context impl;
};

// ---------------------------------------------------------------------------
// Exception boundary
// ---------------------------------------------------------------------------
template <class F>
static int guard(F &&f) {
    try {
        f();
        set_last_error("");
        return KRT_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return KRT_ERR_BUDGET;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return KRT_ERR_INVALID;
    }
}

extern "C" {

const char *krt_last_error(void) { return g_last_error.c_str(); }

int krt_create(int device, krt_context **out) {
    return guard([&] {
        KRT_REQUIRE(out != nullptr, KRT_ERR_INVALID, "out is null");
        int n = 0;
        KRT_HIP(hipGetDeviceCount(&n));
        KRT_REQUIRE(device >= 0 && device < n, KRT_ERR_INVALID, "no such HIP device");
        KRT_HIP(hipSetDevice(device));
        auto c = std::make_unique<krt_context>();
        c->impl.device = device;
        c->impl.noalloc_spill = noalloc_scratch();  // once per device, with the device current
        KRT_HIP(hipStreamCreate(&c->impl.stream));
        KRT_HIP(hipStreamCreate(&c->impl.side));
        c->impl.counters.alloc(sizeof(unsigned long long) * 16);
        *out = c.release();
    });
}

void krt_destroy(krt_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->impl.device);
    (void)hipStreamSynchronize(ctx->impl.stream);
    if (ctx->impl.side) (void)hipStreamSynchronize(ctx->impl.side);
    ctx->impl.scenes.clear();
    ctx->impl.counters.release();
    ctx->impl.spill.release();
    for (auto &e : ctx->impl.timer)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->impl.stream);
    if (ctx->impl.side) (void)hipStreamDestroy(ctx->impl.side);
    delete ctx;
}

static void upload(Context &C, DevBuf &b, const void *src, size_t bytes) {
    b.alloc(bytes);
    if (bytes) KRT_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, C.stream));
}

// Material rows: {diffuse, rough}, {specular, refr}, {emission, type}, and
// the texture ids {diffuse | normal << 16, specular | emission << 16,
// roughness | 0xFFFF << 16, any} (0xFFFF: none; mat_tex: 5 per material
// {diffuse, normal, specular, emission, roughness}, -1 none, or NULL).
static std::vector<float> material_rows(const float *materials, uint32_t nm, const int32_t *mat_tex = nullptr) {
    std::vector<float> mt(16 * (size_t)nm, 0.0f);
    for (uint32_t m = 0; m < nm; m++) {
        const float *q = materials + 12 * m;
        float *o = &mt[16 * m];
        o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[9];
        o[4] = q[3]; o[5] = q[4]; o[6] = q[5]; o[7] = q[10];
        o[8] = q[6]; o[9] = q[7]; o[10] = q[8];
        int type = (int)q[11];
        std::memcpy(&o[11], &type, 4);
        uint32_t id[5];
        bool any = false;
        for (int k = 0; k < 5; k++) {
            const int32_t t = mat_tex ? mat_tex[5 * m + k] : -1;
            id[k] = t < 0 ? 0xFFFFu : (uint32_t)t;
            any = any || t >= 0;
        }
        const uint32_t row[4] = {id[0] | id[1] << 16, id[2] | id[3] << 16, id[4] | 0xFFFF0000u, any ? 1u : 0u};
        std::memcpy(&o[12], row, sizeof(row));
    }
    return mt;
}

static bool emissive(const float *e) { return std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) > 0.0f; }

// Triangle light row (armada.cu:117-140) with area and normal precomputed by
// the per-sample expressions of lighting.cuh:102-110.
static void light_row(const float *a, const float *b, const float *c, const float *e, std::vector<float> &lt) {
    float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
    float dd = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    float area = std::sqrt(dd) * 0.5f;
    float inv = 1.0f / std::sqrt(dd);
    float nrm[3] = {cr[0] * inv, cr[1] * inv, cr[2] * inv};
    float row[16] = {a[0], a[1], a[2], area, ab[0], ab[1], ab[2], nrm[0],
                     ac[0], ac[1], ac[2], nrm[1], e[0], e[1], e[2], nrm[2]};
    lt.insert(lt.end(), row, row + 16);
}

static void check_textures(const krt_scene_desc *d, const krt_scene_textures *x) {
    KRT_REQUIRE(x->nv == d->nv && x->nm == d->nm, KRT_ERR_INVALID, "texture arrays do not match the scene");
    KRT_REQUIRE(x->uvs && x->material_textures, KRT_ERR_INVALID, "null texture array");
    KRT_REQUIRE(x->ntex == 0 || (x->texels && x->dims), KRT_ERR_INVALID, "null texel array");
    KRT_REQUIRE(x->ntex < 0xFFFFu, KRT_ERR_INVALID, "at most 65534 textures");
    for (uint32_t t = 0; t < x->ntex; t++)
        KRT_REQUIRE(x->dims[2 * t] >= 1 && x->dims[2 * t + 1] >= 1 && x->dims[2 * t] <= 16384 &&
                        x->dims[2 * t + 1] <= 16384,
                    KRT_ERR_INVALID, "texture sides must be in [1, 16384]");
    for (uint32_t m = 0; m < d->nm; m++)
        for (int k = 0; k < 5; k++) {
            const int32_t t = x->material_textures[5 * m + k];
            KRT_REQUIRE(t >= -1 && t < (int32_t)x->ntex, KRT_ERR_INVALID, "material texture index out of range");
            KRT_REQUIRE(!(k == 1 && t >= 0 && !x->tangents), KRT_ERR_INVALID, "a normal map needs tangents");
        }
}

// Emissive test of the triangle-light list (armada.cu:264-265): a non-zero
// emission or an emission texture.
static bool emissive_mat(const krt_scene_desc *d, const krt_scene_textures *x, uint32_t m) {
    return emissive(d->materials + 12 * m + 6) || (x && x->material_textures[5 * m + 3] >= 0);
}

static void scene_upload(krt_context *ctx, const krt_scene_desc *d, const krt_scene_textures *x,
                     krt_handle *scene) {
    KRT_REQUIRE(ctx && d && scene, KRT_ERR_INVALID, "null argument");
    KRT_REQUIRE(d->nt > 0 && d->nv > 0 && d->nm > 0, KRT_ERR_INVALID, "empty scene");
    KRT_REQUIRE(d->positions && d->normals && d->indices && d->tri_material && d->materials, KRT_ERR_INVALID,
                "null scene array");
    Context &C = ctx->impl;
    KRT_HIP(hipSetDevice(C.device));
    for (uint32_t i = 0; i < 3 * d->nt; i++)
        KRT_REQUIRE(d->indices[i] < d->nv, KRT_ERR_INVALID, "vertex index out of range");
    for (uint32_t t = 0; t < d->nt; t++)
        KRT_REQUIRE(d->tri_material[t] < d->nm, KRT_ERR_INVALID, "material index out of range");
    if (x) check_textures(d, x);
    auto s = std::make_unique<Scene>();
    s->nv = d->nv;
    s->nt = d->nt;
    s->nm = d->nm;
    upload(C, s->pos, d->positions, sizeof(float) * 3 * d->nv);
    for (int a = 0; a < 3; a++) s->bmin[a] = s->bmax[a] = d->positions[a];
    for (uint32_t v = 0; v < d->nv; v++)
        for (int a = 0; a < 3; a++) {
            s->bmin[a] = std::min(s->bmin[a], d->positions[3 * v + a]);
            s->bmax[a] = std::max(s->bmax[a], d->positions[3 * v + a]);
        }
    upload(C, s->nrm, d->normals, sizeof(float) * 3 * d->nv);
    upload(C, s->idx, d->indices, sizeof(uint32_t) * 3 * d->nt);
    upload(C, s->tri_mat, d->tri_material, sizeof(uint32_t) * d->nt);
    std::vector<float> mt = material_rows(d->materials, d->nm, x ? x->material_textures : nullptr);
    upload(C, s->mats, mt.data(), sizeof(float) * mt.size());
    if (x) {
        s->textured = true;
        upload(C, s->uv, x->uvs, sizeof(float) * 2 * d->nv);
        if (x->tangents) upload(C, s->tan, x->tangents, sizeof(float) * 6 * d->nv);
        std::vector<uint32_t> desc(4 * (size_t)std::max(x->ntex, 1u), 0u);
        size_t texels = 0;
        for (uint32_t t = 0; t < x->ntex; t++) {
            desc[4 * t] = (uint32_t)texels;
            desc[4 * t + 1] = x->dims[2 * t];
            desc[4 * t + 2] = x->dims[2 * t + 1];
            texels += (size_t)x->dims[2 * t] * x->dims[2 * t + 1];
        }
        KRT_REQUIRE(texels < (1ull << 32), KRT_ERR_INVALID, "more than 2^32 texels");
        upload(C, s->texels, x->texels, sizeof(uint32_t) * texels);
        upload(C, s->tex_desc, desc.data(), sizeof(uint32_t) * desc.size());
    }
    // triangle lights in triangle order (armada.cu:117-140, emissive test
    // :264-265), with the material's constant emission (armada.cu:143)
    std::vector<float> lt;
    std::vector<int32_t> ltri;
    for (uint32_t t = 0; t < d->nt; t++) {
        if (!emissive_mat(d, x, d->tri_material[t])) continue;
        light_row(d->positions + 3 * d->indices[3 * t], d->positions + 3 * d->indices[3 * t + 1],
                  d->positions + 3 * d->indices[3 * t + 2], d->materials + 12 * d->tri_material[t] + 6, lt);
        ltri.push_back((int32_t)t);
    }
    s->nl = (uint32_t)(lt.size() / 16);
    upload(C, s->lights, lt.data(), sizeof(float) * lt.size());
    upload(C, s->light_tri, ltri.data(), sizeof(int32_t) * ltri.size());
    // area lights of the editor's ReSTIR (one per emissive material, its
    // triangles in triangle order)
    {
        std::map<uint32_t, uint32_t> of_mat;
        std::vector<int32_t> count;
        std::vector<float> em;
        for (uint32_t t = 0; t < d->nt; t++) {
            const uint32_t m = d->tri_material[t];
            const float *e = d->materials + 12 * m + 6;
            if (!emissive(e)) continue;
            auto it = of_mat.find(m);
            if (it == of_mat.end()) {
                it = of_mat.emplace(m, (uint32_t)count.size()).first;
                count.push_back(0);
                em.insert(em.end(), {e[0], e[1], e[2], 0.0f});
            }
            count[it->second]++;
        }
        std::vector<int32_t> range(2 * count.size()), fill(count.size(), 0), tris;
        int32_t run = 0;
        for (size_t l = 0; l < count.size(); l++) {
            range[2 * l] = run;
            range[2 * l + 1] = count[l];
            run += count[l];
        }
        tris.assign((size_t)run, 0);
        for (uint32_t t = 0; t < d->nt; t++) {
            auto it = of_mat.find(d->tri_material[t]);
            if (it == of_mat.end()) continue;
            tris[(size_t)range[2 * it->second] + fill[it->second]++] = (int32_t)t;
        }
        s->nml = (uint32_t)count.size();
        if (s->nml) {
            upload(C, s->ml_range, range.data(), sizeof(int32_t) * range.size());
            upload(C, s->ml_tris, tris.data(), sizeof(int32_t) * tris.size());
            upload(C, s->ml_emission, em.data(), sizeof(float) * em.size());
        }
    }
    KRT_HIP(hipStreamSynchronize(C.stream));
    krt_handle h = C.next_handle++;
    C.scenes[h] = std::move(s);
    *scene = h;
}

int krt_scene_upload(krt_context *ctx, const krt_scene_desc *d, krt_handle *scene) {
    return guard([&] { scene_upload(ctx, d, nullptr, scene); });
}

int krt_scene_upload_textured(krt_context *ctx, const krt_scene_desc *d, const krt_scene_textures *tex,
                              krt_handle *scene) {
    return guard([&] {
        KRT_REQUIRE(tex, KRT_ERR_INVALID, "null argument");
        scene_upload(ctx, d, tex, scene);
    });
}

int krt_scene_upload_instanced(krt_context *ctx, const krt_instanced_desc *d, krt_handle *scene) {
    return guard([&] {
        KRT_REQUIRE(ctx && d && scene, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(d->nmesh > 0 && d->ninst > 0 && d->nm > 0, KRT_ERR_INVALID, "empty instanced scene");
        KRT_REQUIRE(d->positions && d->normals && d->mesh_nv && d->indices && d->mesh_nt && d->instance_mesh &&
                        d->transforms && d->tri_material && d->materials,
                    KRT_ERR_INVALID, "null scene array");
        Context &C = ctx->impl;
        KRT_HIP(hipSetDevice(C.device));
        auto s = std::make_unique<Scene>();
        s->instanced = true;
        s->nmesh = d->nmesh;
        s->ninst = d->ninst;
        s->nm = d->nm;
        uint64_t nv = 0, ntm = 0;
        for (uint32_t m = 0; m < d->nmesh; m++) {
            KRT_REQUIRE(d->mesh_nv[m] > 0 && d->mesh_nt[m] > 0, KRT_ERR_INVALID, "empty mesh");
            s->mesh_vbase.push_back((uint32_t)nv);
            s->mesh_tbase.push_back((uint32_t)ntm);
            s->mesh_nv.push_back(d->mesh_nv[m]);
            s->mesh_nt.push_back(d->mesh_nt[m]);
            for (uint64_t i = 3 * ntm; i < 3 * (ntm + d->mesh_nt[m]); i++)
                KRT_REQUIRE(d->indices[i] < d->mesh_nv[m], KRT_ERR_INVALID, "vertex index out of range");
            nv += d->mesh_nv[m];
            ntm += d->mesh_nt[m];
            KRT_REQUIRE(d->mesh_nt[m] < (1u << 26), KRT_ERR_INVALID, "too many triangles in a mesh");
        }
        KRT_REQUIRE(nv < (1ull << 31) && ntm < (1ull << 31), KRT_ERR_INVALID, "meshes too large");
        uint64_t nt = 0;
        for (uint32_t i = 0; i < d->ninst; i++) {
            KRT_REQUIRE(d->instance_mesh[i] < d->nmesh, KRT_ERR_INVALID, "instance mesh out of range");
            for (int k = 0; k < 12; k++)
                KRT_REQUIRE(std::isfinite(d->transforms[12 * (size_t)i + k]), KRT_ERR_INVALID,
                            "instance transform has non-finite entries");
            float inv[12];
            affine_inverse(d->transforms + 12 * (size_t)i, inv);  // rejects singular transforms
            s->inst_mesh.push_back((int32_t)d->instance_mesh[i]);
            nt += d->mesh_nt[d->instance_mesh[i]];
        }
        KRT_REQUIRE(nt < (1ull << 31), KRT_ERR_INVALID, "too many instanced triangles");
        KRT_REQUIRE(d->ninst < (1u << 26), KRT_ERR_INVALID, "too many instances");
        for (uint64_t t = 0; t < nt; t++)
            KRT_REQUIRE(d->tri_material[t] < d->nm, KRT_ERR_INVALID, "material index out of range");
        s->xform.assign(d->transforms, d->transforms + 12 * (size_t)d->ninst);
        s->nv = (uint32_t)nv;
        s->nt = (uint32_t)nt;
        upload(C, s->pos, d->positions, sizeof(float) * 3 * nv);
        upload(C, s->nrm, d->normals, sizeof(float) * 3 * nv);
        upload(C, s->idx, d->indices, sizeof(uint32_t) * 3 * ntm);
        upload(C, s->tri_mat, d->tri_material, sizeof(uint32_t) * nt);
        std::vector<float> mt = material_rows(d->materials, d->nm);
        upload(C, s->mats, mt.data(), sizeof(float) * mt.size());
        // world-space triangle lights, instance order then triangle order, the
        // vertices through the instance transform (Transform::apply,
        // armada.cu:131-133); vertex bounding box of the world-space scene
        std::vector<float> lt;
        for (int a = 0; a < 3; a++) {
            s->bmin[a] = FLT_MAX;
            s->bmax[a] = -FLT_MAX;
        }
        uint64_t g = 0;
        for (uint32_t i = 0; i < d->ninst; i++) {
            const uint32_t m = d->instance_mesh[i];
            const float *M = d->transforms + 12 * (size_t)i;
            const float *P = d->positions + 3 * (size_t)s->mesh_vbase[m];
            const uint32_t *I = d->indices + 3 * (size_t)s->mesh_tbase[m];
            auto xf = [&](uint32_t v, float *w) {
                const float *p = P + 3 * (size_t)v;
                for (int r = 0; r < 3; r++)
                    w[r] = (M[4 * r] * p[0] + M[4 * r + 1] * p[1]) + (M[4 * r + 2] * p[2] + M[4 * r + 3] * 1.0f);
            };
            for (uint32_t v = 0; v < d->mesh_nv[m]; v++) {
                float w[3];
                xf(v, w);
                for (int a = 0; a < 3; a++) {
                    s->bmin[a] = std::min(s->bmin[a], w[a]);
                    s->bmax[a] = std::max(s->bmax[a], w[a]);
                }
            }
            for (uint32_t t = 0; t < d->mesh_nt[m]; t++, g++) {
                const float *e = d->materials + 12 * d->tri_material[g] + 6;
                if (!emissive(e)) continue;
                float a[3], b[3], c[3];
                xf(I[3 * t], a);
                xf(I[3 * t + 1], b);
                xf(I[3 * t + 2], c);
                light_row(a, b, c, e, lt);
            }
        }
        s->nl = (uint32_t)(lt.size() / 16);
        upload(C, s->lights, lt.data(), sizeof(float) * lt.size());
        KRT_HIP(hipStreamSynchronize(C.stream));
        krt_handle h = C.next_handle++;
        C.scenes[h] = std::move(s);
        *scene = h;
    });
}

int krt_accel_instance_tree(krt_context *ctx, krt_handle scene, float *out, size_t nfloats, uint32_t *nodes) {
    return guard([&] {
        KRT_REQUIRE(ctx && nodes, KRT_ERR_INVALID, "null argument");
        const Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(s.instanced && s.tlas.built, KRT_ERR_STATE, "not a built instanced scene");
        *nodes = s.tlas.nser;
        if (!out) return;
        KRT_REQUIRE(nfloats >= 12 * (size_t)s.tlas.nser, KRT_ERR_INVALID, "output too small");
        KRT_HIP(hipMemcpy(out, s.tlas.ser.p, sizeof(float) * 12 * (size_t)s.tlas.nser, hipMemcpyDeviceToHost));
    });
}

int krt_accel_mesh_serialized(krt_context *ctx, krt_handle scene, uint32_t mesh, float *out, size_t nfloats,
                              uint32_t *nodes) {
    return guard([&] {
        KRT_REQUIRE(ctx && nodes, KRT_ERR_INVALID, "null argument");
        const Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(s.instanced && s.accel.built, KRT_ERR_STATE, "not a built instanced scene");
        KRT_REQUIRE(mesh < s.nmesh, KRT_ERR_INVALID, "mesh index out of range");
        const uint32_t n = 2 * s.mesh_nt[mesh] - 1;
        *nodes = n;
        if (!out) return;
        KRT_REQUIRE(nfloats >= 12 * (size_t)n, KRT_ERR_INVALID, "output too small");
        KRT_HIP(hipMemcpy(out, s.mesh_ser.as<float4>() + 3 * (size_t)s.mesh_ser_base[mesh],
                          sizeof(float) * 12 * (size_t)n, hipMemcpyDeviceToHost));
    });
}

int krt_scene_free(krt_context *ctx, krt_handle scene) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null context");
        KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
        KRT_REQUIRE(ctx->impl.scene(scene).serve_refs.load() == 0, KRT_ERR_STATE,
                    "scene is used by server requests in flight");
        KRT_REQUIRE(ctx->impl.scenes.erase(scene) == 1, KRT_ERR_INVALID, "unknown scene handle");
    });
}

int krt_scene_light_count(krt_context *ctx, krt_handle scene, uint32_t *out) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        *out = ctx->impl.scene(scene).nl;
    });
}

int krt_scene_lights(krt_context *ctx, krt_handle scene, float *out, size_t nfloats) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(nfloats >= 12 * (size_t)s.nl, KRT_ERR_INVALID, "output too small");
        std::vector<float> rows(16 * (size_t)s.nl);
        if (s.nl)
            KRT_HIP(hipMemcpy(rows.data(), s.lights.p, sizeof(float) * rows.size(), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < s.nl; i++) {
            const float *r = &rows[16 * i];
            float v[12] = {r[0], r[1], r[2], r[4], r[5], r[6], r[8], r[9], r[10], r[12], r[13], r[14]};
            std::memcpy(out + 12 * i, v, sizeof(v));
        }
    });
}

int krt_accel_build_layout(krt_context *ctx, krt_handle scene, int leaf_max, int layout) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null context");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(s.serve_refs.load() == 0, KRT_ERR_STATE, "scene is used by server requests in flight");
        accel_build(ctx->impl, s, leaf_max, layout);
    });
}

int krt_scene_update_vertices(krt_context *ctx, krt_handle scene, uint32_t first, uint32_t count,
                              const float *positions, const float *normals) {
    return guard([&] {
        KRT_REQUIRE(ctx && positions, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "vertex updates take flat scenes (krt_scene_upload)");
        KRT_REQUIRE(s.serve_refs.load() == 0, KRT_ERR_STATE, "scene is used by server requests in flight");
        KRT_REQUIRE((uint64_t)first + count <= s.nv, KRT_ERR_INVALID, "vertex range out of the scene");
        for (size_t i = 0; i < 3 * (size_t)count; i++)
            KRT_REQUIRE(std::isfinite(positions[i]) && (!normals || std::isfinite(normals[i])), KRT_ERR_INVALID,
                        "non-finite vertex data");
        if (count == 0) return;
        hipStream_t st = ctx->impl.stream;
        KRT_HIP(hipStreamSynchronize(st));  // frames in flight may still read the old vertices
        KRT_HIP(hipMemcpyAsync(s.pos.as<float>() + 3 * (size_t)first, positions, sizeof(float) * 3 * count,
                               hipMemcpyHostToDevice, st));
        if (normals)
            KRT_HIP(hipMemcpyAsync(s.nrm.as<float>() + 3 * (size_t)first, normals, sizeof(float) * 3 * count,
                                   hipMemcpyHostToDevice, st));
        KRT_HIP(hipStreamSynchronize(st));
        for (uint32_t v = 0; v < count; v++)  // the box grows to hold the moved vertices
            for (int a = 0; a < 3; a++) {
                s.bmin[a] = std::min(s.bmin[a], positions[3 * v + a]);
                s.bmax[a] = std::max(s.bmax[a], positions[3 * v + a]);
            }
        s.stale = true;
    });
}

int krt_accel_refit(krt_context *ctx, krt_handle scene) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null context");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        Scene &s = ctx->impl.scene(scene);
        KRT_REQUIRE(s.serve_refs.load() == 0, KRT_ERR_STATE, "scene is used by server requests in flight");
        accel_refit(ctx->impl, s);
    });
}

int krt_accel_build(krt_context *ctx, krt_handle scene, int leaf_max) {
    return krt_accel_build_layout(ctx, scene, leaf_max, KRT_LAYOUT_SAH);
}

int krt_accel_stats(krt_context *ctx, krt_handle scene, krt_bvh_stats *out) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        const Accel &A = ctx->impl.scene(scene).accel;
        KRT_REQUIRE(A.built, KRT_ERR_STATE, "acceleration structure not built");
        out->triangles = A.nt;
        out->serialized_nodes = A.nser;
        out->traversal_nodes = A.nwide;
        out->depth_bound = A.depth_bound;
        out->build_levels = A.levels;
        out->build_ms = A.build_ms;
        out->refit_ms = A.refit_ms;
        out->layout = A.layout;
        out->binary_nodes = A.nnodes;
        out->node_bytes = 128;
    });
}

int krt_accel_serialized(krt_context *ctx, krt_handle scene, float *out, size_t nfloats) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        const Accel &A = ctx->impl.scene(scene).accel;
        KRT_REQUIRE(A.built, KRT_ERR_STATE, "acceleration structure not built");
        KRT_REQUIRE(!ctx->impl.scene(scene).instanced, KRT_ERR_STATE,
                    "instanced scene: use krt_accel_mesh_serialized / krt_accel_instance_tree");
        KRT_REQUIRE(nfloats >= 12 * (size_t)A.nser, KRT_ERR_INVALID, "output too small");
        KRT_HIP(hipMemcpy(out, A.ser.p, sizeof(float) * 12 * (size_t)A.nser, hipMemcpyDeviceToHost));
    });
}

int krt_accel_leaf_order(krt_context *ctx, krt_handle scene, int32_t *out, size_t n) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        const Accel &A = ctx->impl.scene(scene).accel;
        KRT_REQUIRE(A.built, KRT_ERR_STATE, "acceleration structure not built");
        KRT_REQUIRE(!ctx->impl.scene(scene).instanced, KRT_ERR_STATE, "instanced scene: no single leaf order");
        KRT_REQUIRE(n >= A.nt, KRT_ERR_INVALID, "output too small");
        KRT_HIP(hipMemcpy(out, A.leaf_order.p, sizeof(int32_t) * A.nt, hipMemcpyDeviceToHost));
    });
}

int krt_trace_closest(krt_context *ctx, krt_handle scene, const float *rays, uint32_t n, float *hits) {
    return guard([&] {
        KRT_REQUIRE(ctx && rays && hits, KRT_ERR_INVALID, "null argument");
        Context &C = ctx->impl;
        Scene &s = C.scene(scene);
        DevBuf r(sizeof(float) * 8 * (size_t)n), h(sizeof(float) * 4 * (size_t)n);
        KRT_HIP(hipMemcpyAsync(r.p, rays, r.bytes, hipMemcpyHostToDevice, C.stream));
        trace_closest_dev(C, s, r.as<float>(), n, h.as<float>());
        KRT_HIP(hipMemcpyAsync(hits, h.p, sizeof(float) * 4 * (size_t)n, hipMemcpyDeviceToHost, C.stream));
        KRT_HIP(hipStreamSynchronize(C.stream));
    });
}

int krt_trace_occluded(krt_context *ctx, krt_handle scene, const float *rays, uint32_t n, int32_t *occ) {
    return guard([&] {
        KRT_REQUIRE(ctx && rays && occ, KRT_ERR_INVALID, "null argument");
        Context &C = ctx->impl;
        Scene &s = C.scene(scene);
        DevBuf r(sizeof(float) * 8 * (size_t)n), o(sizeof(int32_t) * (size_t)n);
        KRT_HIP(hipMemcpyAsync(r.p, rays, r.bytes, hipMemcpyHostToDevice, C.stream));
        trace_occluded_dev(C, s, r.as<float>(), n, o.as<int32_t>());
        KRT_HIP(hipMemcpyAsync(occ, o.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, C.stream));
        KRT_HIP(hipStreamSynchronize(C.stream));
    });
}

int krt_trace_closest_device(krt_context *ctx, krt_handle scene, const void *rays_dev, uint32_t n, void *hits_dev) {
    return guard([&] {
        KRT_REQUIRE(ctx && rays_dev && hits_dev, KRT_ERR_INVALID, "null argument");
        trace_closest_dev(ctx->impl, ctx->impl.scene(scene), (const float *)rays_dev, n, (float *)hits_dev);
    });
}

int krt_trace_closest_device_noalloc(krt_context *ctx, krt_handle scene, const void *rays_dev, uint32_t n,
                                     void *hits_dev) {
    return guard([&] {
        KRT_REQUIRE(ctx && rays_dev && hits_dev, KRT_ERR_INVALID, "null argument");
        trace_closest_dev_noalloc(ctx->impl, ctx->impl.scene(scene), (const float *)rays_dev, n, (float *)hits_dev);
    });
}

int krt_camera_rays(krt_context *ctx, const krt_camera *cam, uint32_t width, uint32_t height, int jitter,
                    int sample, uint32_t user_seed, float *rays) {
    return guard([&] {
        KRT_REQUIRE(ctx && cam && rays, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(width > 0 && height > 0, KRT_ERR_INVALID, "zero extent");
        Context &C = ctx->impl;
        LaunchInfo li{};
        li.width = width;
        li.height = height;
        std::memcpy(li.center, cam->center, sizeof(li.center));
        std::memcpy(li.U, cam->U, sizeof(li.U));
        std::memcpy(li.V, cam->V, sizeof(li.V));
        std::memcpy(li.W, cam->W, sizeof(li.W));
        li.samples = sample;
        li.user_seed = user_seed;
        DevBuf r(sizeof(float) * 8 * (size_t)width * height);
        camera_rays_dev(C, li, jitter, r.as<float>());
        KRT_HIP(hipMemcpyAsync(rays, r.p, r.bytes, hipMemcpyDeviceToHost, C.stream));
        KRT_HIP(hipStreamSynchronize(C.stream));
    });
}

int krt_copy_probe(krt_context *ctx, size_t bytes, int reps, float *ms) {
    return guard([&] {
        KRT_REQUIRE(ctx && ms, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(bytes >= 16 && reps >= 1, KRT_ERR_INVALID, "bytes >= 16 and reps >= 1 required");
        *ms = copy_probe(ctx->impl, bytes, reps);
    });
}

int krt_timer_begin(krt_context *ctx) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null argument");
        Context &C = ctx->impl;
        for (auto &e : C.timer)
            if (!e) KRT_HIP(hipEventCreate(&e));
        KRT_HIP(hipEventRecord(C.timer[0], C.stream));
    });
}

int krt_timer_end(krt_context *ctx, float *ms) {
    return guard([&] {
        KRT_REQUIRE(ctx && ms, KRT_ERR_INVALID, "null argument");
        Context &C = ctx->impl;
        KRT_REQUIRE(C.timer[0] && C.timer[1], KRT_ERR_STATE, "krt_timer_begin was not called");
        KRT_HIP(hipEventRecord(C.timer[1], C.stream));
        KRT_HIP(hipEventSynchronize(C.timer[1]));
        KRT_HIP(hipEventElapsedTime(ms, C.timer[0], C.timer[1]));
    });
}

int krt_sync(krt_context *ctx) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int krt_buffer_create(krt_context *ctx, size_t bytes, void **dev) {
    return guard([&] {
        KRT_REQUIRE(ctx && dev, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(bytes > 0, KRT_ERR_INVALID, "empty buffer");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        *dev = nullptr;
        KRT_HIP(hipMalloc(dev, bytes));
    });
}

int krt_buffer_destroy(krt_context *ctx, void *dev) {
    return guard([&] {
        KRT_REQUIRE(ctx, KRT_ERR_INVALID, "null argument");
        if (!dev) return;
        KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
        KRT_HIP(hipFree(dev));
    });
}

int krt_buffer_upload(krt_context *ctx, void *dev, const void *host, size_t bytes) {
    return guard([&] {
        KRT_REQUIRE(ctx && dev && host, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->impl.stream));
        KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int krt_buffer_download(krt_context *ctx, void *host, const void *dev, size_t bytes) {
    return guard([&] {
        KRT_REQUIRE(ctx && dev && host, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->impl.stream));
        KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

// ------------------------------------------------------------------- armada
int krt_armada_create(krt_context *ctx, uint32_t width, uint32_t height, krt_armada **out) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(width > 0 && height > 0, KRT_ERR_INVALID, "zero extent");
        KRT_REQUIRE((uint64_t)width * height < (1ull << 31), KRT_ERR_INVALID, "extent too large");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        auto *a = new krt_armada();
        a->ctx = &ctx->impl;
        a->w = width;
        a->h = height;
        size_t bytes = sizeof(float) * 4 * (size_t)width * height;
        try {
            a->color.alloc(bytes);
            a->normal.alloc(bytes);
            a->albedo.alloc(bytes);
            a->position.alloc(bytes);
            a->ctr.alloc(sizeof(unsigned long long) * 16);
            for (DevBuf *b : {&a->color, &a->normal, &a->albedo, &a->position})
                KRT_HIP(hipMemsetAsync(b->p, 0, bytes, ctx->impl.stream));
            KRT_HIP(hipStreamSynchronize(ctx->impl.stream));
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

void krt_armada_destroy(krt_armada *a) {
    if (!a) return;
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;  // attachments keep their own buffers until krt_attachment_destroy
}

int krt_attachment_create(krt_context *ctx, const char *kind, krt_attachment **out) {
    return guard([&] {
        KRT_REQUIRE(ctx && kind && out, KRT_ERR_INVALID, "null argument");
        std::string k(kind);
        KRT_REQUIRE(k == "path_tracer" || k == "restir" || k == "restir_gi" || k == "mamba", KRT_ERR_INVALID,
                    "unknown attachment kind: " + k);
        auto *a = new krt_attachment();
        a->ctx = &ctx->impl;
        a->kind = k;
        if (k == "restir_gi") {  // ReSTIR GI defaults: temporal M-cap 30, 2 spatial passes
            a->ropt.m_cap = 30;
            a->ropt.spatial_passes = 2;
        }
        *out = a;
    });
}

void krt_attachment_destroy(krt_attachment *att) {
    if (!att) return;
    if (att->loaded) {
        (void)hipStreamSynchronize(att->ctx->stream);
        att->unload();
    }
    delete att;
}

int krt_armada_attach(krt_armada *a, const char *name, krt_attachment *att) {
    return guard([&] {
        KRT_REQUIRE(a && name && att, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(att->ctx == a->ctx, KRT_ERR_INVALID, "attachment belongs to another context");
        a->attachments[name] = att;  // armada.cuh:304-311: attach makes it active
        a->active = name;
    });
}

int krt_armada_activate(krt_armada *a, const char *name) {
    return guard([&] {
        KRT_REQUIRE(a && name, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(a->attachments.count(name), KRT_ERR_INVALID, std::string("no attachment named ") + name);
        a->active = name;
    });
}

int krt_armada_set_option(krt_armada *a, const char *field, const krt_option *value) {
    return guard([&] {
        KRT_REQUIRE(a && field && value, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(!a->active.empty(), KRT_ERR_STATE, "no active attachment");
        KRT_REQUIRE(a->attachments.at(a->active)->set_option(field, *value), KRT_ERR_INVALID,
                    std::string("bad value or value type for option ") + field);
    });
}

int krt_armada_get_option(krt_armada *a, const char *field, krt_option *out) {
    return guard([&] {
        KRT_REQUIRE(a && field && out, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(!a->active.empty(), KRT_ERR_STATE, "no active attachment");
        *out = a->attachments.at(a->active)->get_option(field);
    });
}

int krt_armada_set_envmap(krt_armada *a, const float *rgba, uint32_t width, uint32_t height) {
    return guard([&] {
        KRT_REQUIRE(a, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipSetDevice(a->ctx->device));
        if (!rgba || width == 0 || height == 0) {
            KRT_HIP(hipStreamSynchronize(a->ctx->stream));
            a->env.release();
            a->env_w = a->env_h = 0;
            a->env_on = false;
            return;
        }
        KRT_REQUIRE((uint64_t)width * height < (1ull << 28), KRT_ERR_INVALID, "environment map too large");
        for (size_t i = 0; i < 4 * (size_t)width * height; i++)
            KRT_REQUIRE(std::isfinite(rgba[i]), KRT_ERR_INVALID, "environment map has non-finite texels");
        const size_t bytes = sizeof(float) * 4 * (size_t)width * height;
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));  // frames in flight may still read the old map
        if (a->env.bytes < bytes) a->env.alloc(bytes);
        KRT_HIP(hipMemcpyAsync(a->env.p, rgba, bytes, hipMemcpyHostToDevice, a->ctx->stream));
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
        a->env_w = width;
        a->env_h = height;
        a->env_on = true;
    });
}

int krt_armada_set_envmap_enabled(krt_armada *a, int enabled) {
    return guard([&] {
        KRT_REQUIRE(a, KRT_ERR_INVALID, "null argument");
        a->env_on = enabled != 0;
    });
}

int krt_armada_set_depth(krt_armada *a, int depth) {
    return guard([&] {
        KRT_REQUIRE(a, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(depth >= 0 && depth <= 64, KRT_ERR_INVALID, "depth must be in [0, 64]");
        a->max_depth = depth;
    });
}

int krt_armada_render(krt_armada *a, krt_handle scene, const krt_camera *cam, int accumulate) {
    return guard([&] {
        KRT_REQUIRE(a && cam, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(!a->active.empty(), KRT_ERR_STATE, "no active attachment");
        for (int k = 0; k < 3; k++)
            KRT_REQUIRE(std::isfinite(cam->center[k]) && std::isfinite(cam->U[k]) && std::isfinite(cam->V[k]) &&
                            std::isfinite(cam->W[k]),
                        KRT_ERR_INVALID, "camera has non-finite components");
        Context &C = *a->ctx;
        KRT_HIP(hipSetDevice(C.device));
        Scene &s = C.scene(scene);
        KRT_REQUIRE(s.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
        KRT_REQUIRE(!s.stale, KRT_ERR_STATE, "vertices moved: call krt_accel_refit or krt_accel_build first");
        // attachment switch (armada.cu:747-753)
        if (a->active != a->previous) {
            if (!a->previous.empty() && a->attachments.count(a->previous)) {
                KRT_HIP(hipStreamSynchronize(C.stream));
                a->attachments[a->previous]->unload();
            }
            a->previous = a->active;
        }
        krt_attachment *att = a->attachments.at(a->previous);
        if (!att->loaded || att->w != a->w || att->h != a->h) att->load(a->w, a->h);
        // armada.cu:771-772: restart accumulation on a new acceleration structure
        if (!accumulate || scene != a->last_scene || s.version != a->last_version) a->samples = 0;
        a->last_scene = scene;
        a->last_version = s.version;
        LaunchInfo li{};
        li.width = a->w;
        li.height = a->h;
        std::memcpy(li.center, cam->center, sizeof(li.center));
        std::memcpy(li.U, cam->U, sizeof(li.U));
        std::memcpy(li.V, cam->V, sizeof(li.V));
        std::memcpy(li.W, cam->W, sizeof(li.W));
        std::memcpy(li.pv, cam->pv, sizeof(li.pv));
        li.max_depth = a->max_depth;
        li.samples = a->samples;
        li.accumulate = 1;  // ArmadaRTX launches always accumulate (armada.cu:58); samples==0 restarts
        li.user_seed = att->user_seed;
        li.jitter = att->jitter;
        li.color = a->color.as<float4>();
        li.normal = a->normal.as<float4>();
        li.albedo = a->albedo.as<float4>();
        li.position = a->position.as<float4>();
        if (a->env_on && a->env_w > 0) {  // armada.cuh:76-77 (has_environment_map)
            li.env = a->env.as<float4>();
            li.env_w = a->env_w;
            li.env_h = a->env_h;
        }
        KRT_HIP(hipMemsetAsync(a->ctr.p, 0, a->ctr.bytes, C.stream));
        PassTimes times;
        if (att->kind == "path_tracer" && att->copt.enabled) {
            cache_pt_render(C, s, li, att->copt, att->cst, a->ctr.as<unsigned long long>(), times);
        } else if (att->kind == "path_tracer") {
            pt_render(C, s, li, a->ctr.as<unsigned long long>(), times,
                      a->f64 ? a->acc64.as<double>() : nullptr);
            if (a->f64) a->f64_frames++;
        }
        else if (att->kind == "mamba")
            mamba_render(C, s, li, att->mopt, att->mst, times);
        else if (att->kind == "restir")
            restir_render(C, s, li, att->ropt, att->rst, a->ctr.as<unsigned long long>(), times, &att->copt,
                          &att->cst);
        else gi_render(C, s, li, att->ropt, att->gst, a->ctr.as<unsigned long long>(), times);
        unsigned long long c[16] = {0};
        KRT_HIP(hipMemcpyAsync(c, a->ctr.p, sizeof(c), hipMemcpyDeviceToHost, C.stream));
        KRT_HIP(hipStreamSynchronize(C.stream));
        a->stats.closest_rays = c[0];
        a->stats.shadow_rays = c[1] + c[2] + c[3];
        a->stats.pass_rays[0] = c[0] + c[1];
        a->stats.pass_rays[1] = c[2];
        a->stats.pass_rays[2] = c[3];
        a->stats.pass_rays[3] = 0;
        a->stats.node_visits = c[4] + c[5] + c[6];
        a->stats.tri_tests = c[7] + c[8] + c[9];
        for (int k = 0; k < 3; k++) {
            a->stats.pass_nodes[k] = c[4 + k];
            a->stats.pass_tris[k] = c[7 + k];
        }
        a->stats.pass_nodes[3] = a->stats.pass_tris[3] = 0;
        a->stats.samples = (uint64_t)a->w * a->h;
        float tot = 0.0f;
        for (int k = 0; k < 8; k++) {
            a->stats.kernel_ms[k] = times.ms[k];
            tot += times.ms[k];
        }
        a->stats.gpu_ms = times.frame >= 0.0f ? times.frame : tot;
        a->samples++;  // armada.cu:779
    });
}

int krt_armada_buffers(krt_armada *a, krt_buffers *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        out->color = a->color.p;
        out->normal = a->normal.p;
        out->albedo = a->albedo.p;
        out->position = a->position.p;
    });
}

static DevBuf &aov(krt_armada *a, int which) {
    switch (which) {
        case 0: return a->color;
        case 1: return a->normal;
        case 2: return a->albedo;
        case 3: return a->position;
    }
    throw Error(KRT_ERR_INVALID, "buffer index must be 0..3");
}

int krt_armada_readback(krt_armada *a, int which, float *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        DevBuf &b = aov(a, which);
        KRT_HIP(hipMemcpyAsync(out, b.p, b.bytes, hipMemcpyDeviceToHost, a->ctx->stream));
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
    });
}

int krt_armada_upload(krt_armada *a, int which, const float *in) {
    return guard([&] {
        KRT_REQUIRE(a && in, KRT_ERR_INVALID, "null argument");
        DevBuf &b = aov(a, which);
        KRT_HIP(hipMemcpyAsync(b.p, in, b.bytes, hipMemcpyHostToDevice, a->ctx->stream));
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
    });
}

int krt_armada_read_reservoirs(krt_armada *a, int which, float *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(!a->previous.empty(), KRT_ERR_STATE, "nothing rendered yet");
        krt_attachment *att = a->attachments.at(a->previous);
        KRT_REQUIRE(att->kind != "path_tracer" && att->loaded, KRT_ERR_STATE, "active attachment has no reservoirs");
        if (att->kind == "restir") restir_read(*a->ctx, att->rst, which, out);
        else if (att->kind == "mamba") mamba_read(*a->ctx, att->mst, which, out);
        else gi_read(*a->ctx, att->gst, which, out);
    });
}

int krt_armada_set_f64_accumulation(krt_armada *a, int enable) {
    return guard([&] {
        KRT_REQUIRE(a, KRT_ERR_INVALID, "null argument");
        a->f64 = enable != 0;
        a->f64_frames = 0;
        if (a->f64) {
            size_t bytes = sizeof(double) * 6 * (size_t)a->w * a->h;  // sums, then sums of squares
            if (a->acc64.bytes < bytes) a->acc64.alloc(bytes);
            KRT_HIP(hipMemsetAsync(a->acc64.p, 0, bytes, a->ctx->stream));
            KRT_HIP(hipStreamSynchronize(a->ctx->stream));
        } else {
            a->acc64.release();
        }
    });
}

int krt_armada_read_f64(krt_armada *a, double *out, uint64_t *frames) {
    return guard([&] {
        KRT_REQUIRE(a && out && frames, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(a->f64, KRT_ERR_STATE, "FP64 accumulation is not enabled");
        KRT_HIP(hipMemcpyAsync(out, a->acc64.p, sizeof(double) * 3 * (size_t)a->w * a->h, hipMemcpyDeviceToHost,
                               a->ctx->stream));
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
        *frames = a->f64_frames;
    });
}

int krt_armada_read_f64_squares(krt_armada *a, double *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(a->f64, KRT_ERR_STATE, "FP64 accumulation is not enabled");
        const size_t n3 = 3 * (size_t)a->w * a->h;
        KRT_HIP(hipMemcpyAsync(out, a->acc64.as<double>() + n3, sizeof(double) * n3, hipMemcpyDeviceToHost,
                               a->ctx->stream));
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
    });
}

int krt_hdr_to_ldr(krt_context *ctx, const void *pixels_dev, void *target_dev, uint32_t width, uint32_t height,
                   int tonemapping) {
    return guard([&] {
        KRT_REQUIRE(ctx && ((pixels_dev && target_dev) || width == 0 || height == 0), KRT_ERR_INVALID,
                    "null argument");
        KRT_HIP(hipSetDevice(ctx->impl.device));
        hdr_to_ldr(ctx->impl, (const float4 *)pixels_dev, (uint32_t *)target_dev, width, height, tonemapping);
    });
}

int krt_armada_read_ldr(krt_armada *a, int tonemapping, uint32_t *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        Context &C = *a->ctx;
        KRT_HIP(hipSetDevice(C.device));
        const size_t n = (size_t)a->w * a->h;
        if (a->ldr.bytes < sizeof(uint32_t) * n) a->ldr.alloc(sizeof(uint32_t) * n);
        hdr_to_ldr(C, a->color.as<float4>(), a->ldr.as<uint32_t>(), a->w, a->h, tonemapping);
        KRT_HIP(hipMemcpyAsync(out, a->ldr.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, C.stream));
        KRT_HIP(hipStreamSynchronize(C.stream));
    });
}

int krt_write_png(const char *path, const uint32_t *rgba, uint32_t width, uint32_t height) {
    return guard([&] {
        KRT_REQUIRE(path && rgba, KRT_ERR_INVALID, "null argument");
        write_png(path, rgba, width, height);
    });
}

int krt_read_png(const char *path, uint32_t *width, uint32_t *height, uint32_t *rgba, size_t capacity) {
    return guard([&] {
        KRT_REQUIRE(path && width && height, KRT_ERR_INVALID, "null argument");
        if (!rgba) {
            read_png_size(path, *width, *height);
            return;
        }
        std::vector<uint32_t> px;
        uint32_t w = 0, h = 0;
        read_png(path, px, w, h);
        *width = w;
        *height = h;
        KRT_REQUIRE(capacity >= px.size(), KRT_ERR_INVALID, "rgba holds fewer than width x height texels");
        std::memcpy(rgba, px.data(), sizeof(uint32_t) * px.size());
    });
}

int krt_armada_write_png(krt_armada *a, int tonemapping, const char *path) {
    return guard([&] {
        KRT_REQUIRE(a && path, KRT_ERR_INVALID, "null argument");
        std::vector<uint32_t> px((size_t)a->w * a->h);
        const int rc = krt_armada_read_ldr(a, tonemapping, px.data());
        if (rc != KRT_OK) throw Error(rc, krt_last_error());
        write_png(path, px.data(), a->w, a->h);
    });
}

int krt_armada_stats(krt_armada *a, krt_frame_stats *out) {
    return guard([&] {
        KRT_REQUIRE(a && out, KRT_ERR_INVALID, "null argument");
        *out = a->stats;
    });
}

int krt_armada_sync(krt_armada *a) {
    return guard([&] {
        KRT_REQUIRE(a, KRT_ERR_INVALID, "null argument");
        KRT_HIP(hipStreamSynchronize(a->ctx->stream));
    });
}

krt_attachment_ret krt_load_attachment(krt_context *ctx, const char *kind) {
    krt_attachment_ret r{nullptr, nullptr};
    if (krt_attachment_create(ctx, kind, &r.ptr) == KRT_OK) r.name = r.ptr->kind.c_str();
    return r;
}

}  // extern "C"

// --------------------------------------------------------- radiance cache

namespace {

struct HostArray {  // device copy of a host array for the host-pointer entry points
    DevBuf buf;
    HostArray(Context &ctx, const float *src, size_t count) : buf(sizeof(float) * count) {
        KRT_HIP(hipMemcpyAsync(buf.p, src, sizeof(float) * count, hipMemcpyHostToDevice, ctx.stream));
    }
};

}  // namespace

void krt_mlp_default_config(krt_mlp_config *cfg) {
    if (!cfg) return;
    MlpConfig d;
    MlpEncoding e;
    cfg->lr = d.lr;
    cfg->beta1 = d.b1;
    cfg->beta2 = d.b2;
    cfg->eps = d.eps;
    cfg->loss_scale = d.loss_scale;
    cfg->growth_interval = d.growth_interval;
    for (int a = 0; a < 3; a++) {
        cfg->pos_min[a] = e.pmin[a];
        cfg->pos_scale[a] = e.pscale[a];
    }
}

int krt_mlp_create(krt_context *ctx, const krt_mlp_config *cfg, uint64_t seed, krt_mlp **out) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        krt_mlp_config c;
        krt_mlp_default_config(&c);
        if (cfg) c = *cfg;
        KRT_REQUIRE(c.lr > 0.0f && c.beta1 >= 0.0f && c.beta1 < 1.0f && c.beta2 >= 0.0f && c.beta2 < 1.0f &&
                        c.eps >= 0.0f && c.loss_scale > 0.0f && c.growth_interval > 0,
                    KRT_ERR_INVALID, "invalid optimiser configuration");
        auto m = std::make_unique<krt_mlp>();
        m->ctx = &ctx->impl;
        m->owned = std::make_unique<Mlp>();
        m->net = m->owned.get();
        m->net->cfg = MlpConfig{c.lr, c.beta1, c.beta2, c.eps, c.loss_scale, c.growth_interval};
        for (int a = 0; a < 3; a++) {
            m->net->enc.pmin[a] = c.pos_min[a];
            m->net->enc.pscale[a] = c.pos_scale[a];
        }
        mlp_init(*m->ctx, *m->net, m->net->cfg, seed);
        *out = m.release();
    });
}

void krt_mlp_destroy(krt_mlp *net) {
    if (!net || !net->owned) return;
    (void)hipStreamSynchronize(net->ctx->stream);
    delete net;
}

int krt_attachment_cache(krt_attachment *att, krt_mlp **out) {
    return guard([&] {
        KRT_REQUIRE(att && out, KRT_ERR_INVALID, "null argument");
        *out = nullptr;
        if (!att->cst.ready) return;
        if (!att->cache_handle) {
            att->cache_handle = std::make_unique<krt_mlp>();
            att->cache_handle->ctx = att->ctx;
            att->cache_handle->net = &att->cst.net;
        }
        *out = att->cache_handle.get();
    });
}

int krt_mlp_get_params(krt_mlp *net, float *params) {
    return guard([&] {
        KRT_REQUIRE(net && params, KRT_ERR_INVALID, "null argument");
        mlp_get_params(*net->ctx, *net->net, params);
    });
}

int krt_mlp_set_params(krt_mlp *net, const float *params) {
    return guard([&] {
        KRT_REQUIRE(net && params, KRT_ERR_INVALID, "null argument");
        mlp_set_params(*net->ctx, *net->net, params);
    });
}

int krt_mlp_set_encoding(krt_mlp *net, const float pos_min[3], const float pos_scale[3]) {
    return guard([&] {
        KRT_REQUIRE(net && pos_min && pos_scale, KRT_ERR_INVALID, "null argument");
        for (int a = 0; a < 3; a++) {
            net->net->enc.pmin[a] = pos_min[a];
            net->net->enc.pscale[a] = pos_scale[a];
        }
    });
}

int krt_mlp_infer_device(krt_mlp *net, const float *raw_dev, uint32_t n, float *out_dev) {
    return guard([&] {
        KRT_REQUIRE(net && (n == 0 || (raw_dev && out_dev)), KRT_ERR_INVALID, "null argument");
        mlp_infer_dev(*net->ctx, *net->net, raw_dev, n, out_dev);
    });
}

int krt_mlp_infer(krt_mlp *net, const float *raw, uint32_t n, float *out) {
    return guard([&] {
        KRT_REQUIRE(net && (n == 0 || (raw && out)), KRT_ERR_INVALID, "null argument");
        if (n == 0) return;
        Context &ctx = *net->ctx;
        HostArray in(ctx, raw, (size_t)kMlpRaw * n);
        DevBuf o(sizeof(float) * kMlpOut * (size_t)n);
        mlp_infer_dev(ctx, *net->net, in.buf.as<float>(), n, o.as<float>());
        KRT_HIP(hipMemcpyAsync(out, o.p, sizeof(float) * kMlpOut * (size_t)n, hipMemcpyDeviceToHost, ctx.stream));
        KRT_HIP(hipStreamSynchronize(ctx.stream));
    });
}

static void read_state(Context &ctx, Mlp &net, MlpDevState &st) {
    KRT_HIP(hipMemcpyAsync(&st, net.state.p, sizeof(st), hipMemcpyDeviceToHost, ctx.stream));
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

int krt_mlp_train_step_device(krt_mlp *net, const float *raw_dev, const float *target_dev, uint32_t n) {
    return guard([&] {
        KRT_REQUIRE(net && raw_dev && target_dev, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(n > 0, KRT_ERR_INVALID, "empty training batch");
        mlp_gradient_dev(*net->ctx, *net->net, raw_dev, target_dev, n);
        mlp_apply_dev(*net->ctx, *net->net);
    });
}

int krt_mlp_train_step(krt_mlp *net, const float *raw, const float *target, uint32_t n, float *loss) {
    return guard([&] {
        KRT_REQUIRE(net && raw && target, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(n > 0, KRT_ERR_INVALID, "empty training batch");
        Context &ctx = *net->ctx;
        HostArray in(ctx, raw, (size_t)kMlpRaw * n), tg(ctx, target, (size_t)kMlpOut * n);
        mlp_gradient_dev(ctx, *net->net, in.buf.as<float>(), tg.buf.as<float>(), n);
        mlp_apply_dev(ctx, *net->net);
        MlpDevState st;
        read_state(ctx, *net->net, st);
        if (loss) *loss = st.loss;
    });
}

int krt_mlp_gradient(krt_mlp *net, const float *raw, const float *target, uint32_t n, float *grad, float *loss) {
    return guard([&] {
        KRT_REQUIRE(net && raw && target && grad, KRT_ERR_INVALID, "null argument");
        KRT_REQUIRE(n > 0, KRT_ERR_INVALID, "empty training batch");
        Context &ctx = *net->ctx;
        HostArray in(ctx, raw, (size_t)kMlpRaw * n), tg(ctx, target, (size_t)kMlpOut * n);
        mlp_gradient_dev(ctx, *net->net, in.buf.as<float>(), tg.buf.as<float>(), n);
        KRT_HIP(hipMemcpyAsync(grad, net->net->grad.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                               ctx.stream));
        mlp_clear_flag(ctx, *net->net);
        MlpDevState st;
        read_state(ctx, *net->net, st);
        if (loss) *loss = st.loss;
    });
}

int krt_mlp_apply_gradient(krt_mlp *net, const float *grad) {
    return guard([&] {
        KRT_REQUIRE(net && grad, KRT_ERR_INVALID, "null argument");
        mlp_apply_host_gradient(*net->ctx, *net->net, grad);
        KRT_HIP(hipStreamSynchronize(net->ctx->stream));
    });
}

int krt_mlp_get_state(krt_mlp *net, krt_mlp_state *state) {
    return guard([&] {
        KRT_REQUIRE(net && state, KRT_ERR_INVALID, "null argument");
        MlpDevState st;
        read_state(*net->ctx, *net->net, st);
        static_assert(sizeof(st) == sizeof(krt_mlp_state), "state layout");
        std::memcpy(state, &st, sizeof(st));
    });
}

// ---------------------------------------------------------------- serving
struct krt_server {
    krt_context *ctx = nullptr;
    Server *impl = nullptr;
};

void krt_server_default_config(krt_server_config *cfg) {
    if (!cfg) return;
    ServeConfig d;
    cfg->hbm_budget = d.hbm_budget;
    cfg->max_active = d.max_active;
    cfg->use_graph = d.use_graph;
    cfg->max_queue_ms = d.max_queue_ms;
    cfg->step_pixels = d.step_pixels;
    cfg->wave_pixels = d.wave_pixels;
}

int krt_server_create(krt_context *ctx, const krt_server_config *cfg, krt_server **out) {
    return guard([&] {
        KRT_REQUIRE(ctx && out, KRT_ERR_INVALID, "null argument");
        krt_server_config c;
        krt_server_default_config(&c);
        if (cfg) c = *cfg;
        ServeConfig sc;
        sc.hbm_budget = c.hbm_budget;
        sc.max_active = c.max_active;
        sc.use_graph = c.use_graph;
        sc.max_queue_ms = c.max_queue_ms;
        KRT_REQUIRE(c.step_pixels >= 1, KRT_ERR_INVALID, "step_pixels must be >= 1");
        sc.step_pixels = c.step_pixels;
        sc.wave_pixels = c.wave_pixels;
        auto srv = std::make_unique<krt_server>();
        srv->ctx = ctx;
        srv->impl = serve_create(ctx->impl, sc);
        *out = srv.release();
    });
}

void krt_server_destroy(krt_server *srv) {
    if (!srv) return;
    serve_destroy(srv->impl);
    delete srv;
}

int krt_server_submit(krt_server *srv, const krt_request *req, uint64_t *id) {
    return guard([&] {
        KRT_REQUIRE(srv && req && id, KRT_ERR_INVALID, "null argument");
        Scene &s = srv->ctx->impl.scene(req->scene);
        KRT_REQUIRE(!s.instanced, KRT_ERR_STATE, "the server renders flat scenes (krt_scene_upload)");
        KRT_HIP(hipStreamSynchronize(srv->ctx->impl.stream));  // scene uploads and builds are complete
        RestirOptions ro;
        if (req->kind == KRT_REQUEST_RESTIR) {
            ro.candidates = req->restir.candidates;
            ro.m_cap = req->restir.m_cap;
            ro.spatial_passes = req->restir.spatial_passes;
            ro.neighbours = req->restir.neighbours;
            ro.radius = req->restir.radius;
        }
        *id = serve_submit(srv->impl, s, req->camera, req->width, req->height, req->spp, req->max_depth, req->seed,
                           req->kind, ro);
    });
}

int krt_server_wait(krt_server *srv, uint64_t id, float *color, double *latency_ms) {
    int rc = KRT_OK;
    int g = guard([&] {
        KRT_REQUIRE(srv, KRT_ERR_INVALID, "null argument");
        rc = serve_wait(srv->impl, id, color, latency_ms, nullptr, nullptr);
        if (rc != KRT_OK) throw Error(rc, "request rejected (queueing deadline)");
    });
    return g;
}

int krt_server_drain(krt_server *srv) {
    return guard([&] {
        KRT_REQUIRE(srv, KRT_ERR_INVALID, "null argument");
        serve_drain(srv->impl);
    });
}

int krt_server_get_stats(krt_server *srv, krt_server_stats *out) {
    return guard([&] {
        KRT_REQUIRE(srv && out, KRT_ERR_INVALID, "null argument");
        serve_stats(srv->impl, *out);
    });
}

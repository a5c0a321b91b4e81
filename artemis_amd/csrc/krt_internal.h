// Host-side internals shared by the krt translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/krt.h"

namespace krt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define KRT_HIP(call)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw ::krt::Error(KRT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define KRT_REQUIRE(cond, code, msg)                                                                 \
    do {                                                                                             \
        if (!(cond)) throw ::krt::Error((code), (msg));                                              \
    } while (0)

// Device allocation owned by RAII.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        KRT_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

// Acceleration structure of one scene.
struct Accel {
    bool built = false;
    uint32_t nt = 0;
    uint32_t nser = 0;        // 2N-1 serialised nodes
    uint32_t nnodes = 0;      // traversal nodes
    uint32_t depth_bound = 0;
    uint32_t levels = 0;
    float build_ms = 0.0f;
    float refit_ms = 0.0f;     // last krt_accel_refit
    uint32_t bin_depth = 0;    // stack bound of the binary traversal tree
    int layout = 0;
    DevBuf ser;        // float4 x 3 per node, bvh.cpp layout
    DevBuf leaf_order; // int per rank
    DevBuf tri;        // float4 x 3 per triangle in rank order: {v1, prim}, {e1}, {e2}
    DevBuf nodes;      // float4 x 4 per binary traversal node (collapse input)
    DevBuf wide;       // float4 x 8 per 4-wide traversal node (what the kernels walk)
    uint32_t nwide = 0;
    // refit schedule of the bvh.cpp tree (made at the first refit after a
    // build): internal nodes grouped by depth, and each depth's offset
    DevBuf ser_by_depth;
    std::vector<uint32_t> ser_depth_off;
};

struct Scene {
    uint32_t nv = 0, nt = 0, nm = 0, nl = 0;
    DevBuf pos, nrm;   // float x 3 per vertex
    DevBuf idx;        // uint x 3 per triangle
    DevBuf tri_mat;    // uint per triangle
    DevBuf mats;       // float4 x 4 per material
    DevBuf lights;     // float4 x 4 per light
    DevBuf light_tri;  // int per light: its triangle (flat scenes)
    // material textures (krt_scene_upload_textured): float2 texture
    // coordinates and 6 floats {tangent, bitangent} per vertex, RGBA8 texels
    // of every texture back to back and a {offset, width, height, 0} uint4
    // per texture
    bool textured = false;
    DevBuf uv, tan, texels, tex_desc;
    bool stale = false;  // vertices moved since the last build / refit
    // the editor's area lights (Mamba): emissive triangles grouped by
    // material in order of their first triangle; per light {first, count}
    // into ml_tris, and its emission
    uint32_t nml = 0;
    DevBuf ml_range, ml_tris, ml_emission;
    Accel accel;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};  // vertex bounding box
    uint64_t version = 0;  // bumped on every rebuild (resets accumulation)
    // server requests that hold this scene (queued or rendering): while any
    // do, the scene may be neither freed nor rebuilt
    std::atomic<uint32_t> serve_refs{0};
    // Two-level scenes (krt_scene_upload_instanced).  pos / nrm / idx then hold
    // the base meshes back to back in their own space (indices local to their
    // mesh); tri_mat is per global triangle (instance i's triangles follow
    // those of instances 0..i-1); accel.wide / accel.tri hold every mesh's
    // 4-wide tree and packed triangles back to back; tlas is the tree over the
    // instances and inst_rec their records in its traversal order.
    bool instanced = false;
    uint32_t nmesh = 0, ninst = 0;
    std::vector<uint32_t> mesh_nv, mesh_nt, mesh_vbase, mesh_tbase;
    std::vector<int32_t> inst_mesh;
    std::vector<float> xform;                // 12 floats per instance (row-major 3x4)
    std::vector<uint32_t> mesh_ser_base;     // first serialised node of each mesh in mesh_ser
    DevBuf mesh_ser;                         // every mesh's bvh.cpp tree, back to back
    Accel tlas;
    DevBuf inst_rec;
    size_t bytes() const {
        return pos.bytes + nrm.bytes + idx.bytes + tri_mat.bytes + mats.bytes + lights.bytes + uv.bytes + tan.bytes +
               texels.bytes + tex_desc.bytes + accel.ser.bytes +
               accel.leaf_order.bytes + accel.tri.bytes + accel.nodes.bytes + accel.wide.bytes + mesh_ser.bytes +
               tlas.ser.bytes + tlas.leaf_order.bytes + tlas.nodes.bytes + tlas.wide.bytes + inst_rec.bytes;
    }
};

// Material textures on the device: RGBA8 texels of every texture back to
// back and a {first texel, width, height, 0} descriptor per texture.
struct TexSet {
    const uint32_t *texels;
    const uint4 *desc;
};

// Device view of a scene passed to kernels by value.
struct SceneView {
    const float *pos;
    const float *nrm;
    const uint32_t *idx;
    const uint32_t *tri_mat;
    const float4 *mats;
    const float4 *lights;
    uint32_t nl;
    const float4 *tri;
    const float4 *nodes;  // binary traversal tree
    const float4 *wide;   // its 4-wide collapse (instanced scenes: every mesh's, concatenated)
    const float4 *tlas;   // instanced scenes: 4-wide instance tree
    const float4 *inst;   // instanced scenes: instance records (krt_traverse.h)
    // material textures (flat scenes): per-vertex texture coordinates and
    // {tangent, bitangent} (nullptr when absent), the texels and descriptors
    const float2 *uv;
    const float *tan;
    TexSet tex;
};

SceneView make_view(const Scene &s);

// Per-context scratch of the wavefront path tracer (grown on demand; every
// launch sequence uses it from start to end on the context stream).
struct WaveScratch {
    DevBuf hits, inst, state, hist, occ;
    DevBuf v0;  // ReSTIR GI: primary-hit state (4 float4 per pixel)
    DevBuf lec;  // path tracer with the radiance cache: emission at the cache vertex (float4 per pixel)
    DevBuf rays[2], counts[2], prefix[2], dense[2];  // closest-ray list, shadow-ray list
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream for work that overlaps the main stream's (radiance-cache inference)
    std::map<krt_handle, std::unique_ptr<Scene>> scenes;
    krt_handle next_handle = 1;
    DevBuf counters;  // scratch counters (unsigned long long x 16)
    DevBuf spill;     // traversal stack spill slab (grown on demand)
    WaveScratch wave;
    hipEvent_t timer[2] = {nullptr, nullptr};  // krt_timer_begin / krt_timer_end
    int *noalloc_spill = nullptr;  // this device's address of the file-scope slab of krt_trace.hip
    Scene &scene(krt_handle h) {
        auto it = scenes.find(h);
        KRT_REQUIRE(it != scenes.end(), KRT_ERR_INVALID, "unknown scene handle");
        return *it->second;
    }
};

// PNG file of packed RGBA8 pixels (krt_image.cpp)
void write_png(const std::string &path, const uint32_t *rgba, uint32_t width, uint32_t height);
// PNG file -> packed RGBA8 (r in the low byte), rows in file order
void read_png(const std::string &path, std::vector<uint32_t> &rgba, uint32_t &width, uint32_t &height);
void read_png_size(const std::string &path, uint32_t &width, uint32_t &height);

// Builders / kernels (defined in the .hip units)
void accel_build(Context &ctx, Scene &s, int leaf_max, int layout);
void accel_refit(Context &ctx, Scene &s);
void upload_scene_derived(Context &ctx, Scene &s, const float *mats12, uint32_t nm);
void trace_closest_dev(Context &ctx, Scene &s, const float *rays, uint32_t n, float *hits);
void trace_occluded_dev(Context &ctx, Scene &s, const float *rays, uint32_t n, int32_t *occ);
// The closest-hit query without allocation: the traversal-stack overflow slab
// is a file-scope __device__ array of kNoAllocSpillInts ints (16 MiB) instead
// of the grown-on-demand Context::spill.  noalloc_scratch() returns the
// calling device's address of that array (hipGetSymbolAddress, resolved once
// per device; call it with the device current).
constexpr size_t kNoAllocSpillInts = size_t(4) << 20;
int *noalloc_scratch();
void trace_closest_dev_noalloc(Context &ctx, Scene &s, const float *rays, uint32_t n, float *hits);

// Launch parameters of the rendering kernels (ArmadaLaunchInfo, armada.cuh:43-83)
struct LaunchInfo {
    uint32_t width, height;
    float center[3], U[3], V[3], W[3];
    float pv[16];       // this camera's projection*view (column-major)
    float prev_pv[16];  // previous frame's, for temporal reprojection
    int max_depth;
    int samples;
    int accumulate;
    uint32_t user_seed;
    int jitter;  // path tracer: sub-pixel jitter (1, the reference's) or the fixed -0.5 offset (0)
    float4 *color, *normal, *albedo, *position;
    const float4 *env;  // environment map texels (RGBA, row-major), nullptr when off
    uint32_t env_w, env_h;
};

struct RestirOptions {
    int candidates = 32;
    int m_cap = 200;
    int spatial_passes = 2;
    int neighbours = 5;
    int radius = 30;
};

struct RestirState {
    uint32_t w = 0, h = 0;
    DevBuf res[2];  // 4 float4 planes per buffer (SoA of float4)
    int cur = 0;
    DevBuf gmat;          // uint material id per pixel
    DevBuf guv;           // float2 texture coordinates per pixel (textured scenes)
    DevBuf intermediate;  // float4 per pixel
    DevBuf auxiliary;     // float4 per pixel
    // spatial pass, split around its shadow rays: segmented ray list (2 float4
    // per ray), its segment counts and offsets, the packed list, one occlusion
    // byte per (ray slot, pixel), 1/Z weights per (neighbour, pixel),
    // selection flags per pixel
    DevBuf rays, seg_counts, seg_prefix, dense, occ, zm, flags;
    // wavefront initial pass: hits, hit instances, {seed, levels} per pixel,
    // per-level history and occlusion bytes, the cache vertex's emission, the
    // closest-ray list
    DevBuf wf_hits, wf_inst, wf_state, wf_hist, wf_occ, wf_lec;
    DevBuf cl_rays, cl_counts, cl_prefix, cl_dense;
    float prev_pv[16] = {0};
};

// The editor's G-buffer ReSTIR DI (krt_mamba.hip)
struct MambaOptions {
    int temporal = 1;  // mamba.cu temporal_reuse
    int spatial = 1;   // mamba.cu spatial_reuse
};

struct MambaState {
    uint32_t w = 0, h = 0, user_seed = 0, frame = 0;
    int samples = 0;
    DevBuf R, S, P[2];  // 64 B reservoir records: current, spatial output, previous (double-buffered)
    int pin = 0;
    DevBuf gx, gn, gpm;  // ray-cast G-buffer
    DevBuf guv;          // its texture coordinates (float2 per pixel)
    float prev_pv[16] = {0};
    float prev_cam[12] = {0};
    bool have_prev = false;
};

struct GiState {
    uint32_t w = 0, h = 0;
    DevBuf res[2];  // 5 float4 planes per buffer
    int cur = 0;
    DevBuf gxd[2], gnv[2], gmat[2];  // G-buffer of this and the last frame
    DevBuf guv[2];                   // its texture coordinates (float2 per pixel; textured scenes)
    int gcur = 0;
    DevBuf direct, aux;
    // spatial pass: segmented shadow-ray list, segment counts and offsets,
    // packed list, occlusion bytes, 1/Z weights
    DevBuf rays, seg_counts, seg_prefix, dense, occ, zm;
    float prev_pv[16] = {0};
};

struct PassTimes {
    float ms[8] = {0};
    float frame = -1.0f;  // the frame's GPU span when passes overlap (else the sum of ms)
};

void pt_render(Context &ctx, Scene &s, const LaunchInfo &li, unsigned long long *ctr, PassTimes &times,
               double *acc64);
struct CacheView;
// The path tracer's passes on ctx.stream, untimed.  With a cache view (cv and
// cv->enabled) paths stop at the cache vertex and the pass leaves prefix and
// cv->theta for the composite instead of accumulating colour, and records
// `queried` (when given) on ctx.stream once every query record is written.
void pt_wave(Context &ctx, Scene &s, const LaunchInfo &li, unsigned long long *ctr, double *acc64,
             const CacheView *cv, float4 *prefix, hipEvent_t queried = nullptr);
// Grows each buffer of `ws` to what a path-tracer frame of w x h pixels at
// max_depth needs (dry_run: only counts) and returns the bytes added.
size_t wave_scratch_grow(WaveScratch &ws, uint32_t w, uint32_t h, int max_depth, bool instanced, bool cache,
                         bool dry_run = false);
size_t wave_scratch_bytes(const WaveScratch &ws);
void restir_alloc(Context &ctx, RestirState &st, uint32_t w, uint32_t h, uint32_t user_seed);
void restir_read(Context &ctx, RestirState &st, int which, float *host_out);
void gi_alloc(Context &ctx, GiState &st, uint32_t w, uint32_t h, uint32_t user_seed);
void mamba_alloc(Context &ctx, MambaState &st, uint32_t w, uint32_t h, uint32_t user_seed);
void mamba_read(Context &ctx, MambaState &st, int which, float *host_out);
void gi_read(Context &ctx, GiState &st, int which, float *host_out);
void gi_render(Context &ctx, Scene &s, const LaunchInfo &li, const RestirOptions &opt, GiState &st,
               unsigned long long *ctr, PassTimes &times);
float copy_probe(Context &ctx, size_t bytes, int reps);
void camera_rays_dev(Context &ctx, const LaunchInfo &li, int jitter, float *rays);
// Inverse of a row-major 3x4 affine transform (FP64 cofactors, rounded once);
// throws on a singular transform.
void affine_inverse(const float *M, float *Minv);
// hdr_to_ldr (include/cuda/color.cuh:82-121) on the context stream; no sync.
void hdr_to_ldr(Context &ctx, const float4 *pixels, uint32_t *target, uint32_t width, uint32_t height,
                int tonemapping);



// Radiance-cache MLP (krt_mlp.hip).  Parameters in canonical order: per layer
// W[out][in] row-major, then b[out].
constexpr int kMlpRaw = 13;      // pos, dir, normal, albedo, roughness
constexpr int kMlpEncoded = 62;  // frequency encoding width
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;
constexpr int kMlpLayers = 6;
constexpr int kMlpParams = 20867;

struct MlpEncoding {
    float pmin[3] = {0.0f, 0.0f, 0.0f};
    float pscale[3] = {1.0f, 1.0f, 1.0f};
};

struct MlpConfig {
    float lr = 1e-2f, b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
    float loss_scale = 1024.0f;
    int growth_interval = 2000;
};

// Optimiser state kept on the device so that training steps need no host
// round trip (layout of krt_mlp_state).
struct MlpDevState {
    float scale, b1t, b2t, loss;
    int step, good, skipped, bad;
};

struct Mlp {
    MlpConfig cfg;
    MlpEncoding enc;
    DevBuf params, m, v, grad;  // FP32 master weights and Adam moments
    DevBuf frag_f, frag_b;      // bf16 MFMA fragments of W and W^T
    DevBuf state;               // MlpDevState
    DevBuf partial;             // per-workgroup gradient partials
    DevBuf mid;                 // sums of groups of partials
};

void mlp_init(Context &ctx, Mlp &net, const MlpConfig &cfg, uint64_t seed);
void mlp_set_params(Context &ctx, Mlp &net, const float *host);
void mlp_get_params(Context &ctx, Mlp &net, float *host);
void mlp_infer_dev(Context &ctx, Mlp &net, const float *raw, uint32_t n, float *out);
void mlp_infer_dev(Mlp &net, const float *raw, uint32_t n, float *out, hipStream_t stream);
// n_dev (optional): the batch size on the device, at most n.
void mlp_gradient_dev(Context &ctx, Mlp &net, const float *raw, const float *target, uint32_t n,
                      const uint32_t *n_dev = nullptr);
void mlp_apply_dev(Context &ctx, Mlp &net);
void mlp_apply_host_gradient(Context &ctx, Mlp &net, const float *grad);
void mlp_clear_flag(Context &ctx, Mlp &net);

// Radiance cache attached to the path tracer (krt_cache.hip).
struct CacheOptions {
    int enabled = 0;
    int depth = 1;              // level of the cache vertex (0 = primary hit)
    int train_samples = 65536;  // training paths per frame
    int training = 1;           // 0 freezes the network
    int learn_emission = 0;     // 1: emission at the cache vertex is learned, not added
};

struct CacheState {
    MlpConfig cfg;
    Mlp net;
    bool ready = false;
    uint64_t frame = 0;
    uint32_t npix = 0, slots = 0;
    DevBuf query, pred, prefix, theta;             // per pixel
    DevBuf slot_raw, slot_tgt, slot_ok;            // per training slot
    DevBuf batch_raw, batch_tgt, batch_n;          // compacted training batch
    DevBuf chunk_n;                                // per chunk of slots: valid count, then first batch row
};

void cache_pt_render(Context &ctx, Scene &s, const LaunchInfo &li, const CacheOptions &opt, CacheState &cs,
                     unsigned long long *ctr, PassTimes &times);
// ReSTIR DI; with a cache (copt->enabled) its indirect path stops at the cache vertex.
void mamba_render(Context &ctx, Scene &s, const LaunchInfo &li, const MambaOptions &opt, MambaState &st,
                  PassTimes &times);

// sync = false: everything is queued on ctx.stream and the call returns
// without waiting (no pass times); the radiance cache needs sync = true.
void restir_render(Context &ctx, Scene &s, const LaunchInfo &li, const RestirOptions &opt, RestirState &st,
                   unsigned long long *ctr, PassTimes &times, const CacheOptions *copt = nullptr,
                   CacheState *cs = nullptr, bool sync = true);

}  // namespace krt

/*
 * krt — C ABI of the MI355X ray-traced rendering path.
 *
 * The reference renderer (Kobra, vedavamadathil/artemis) exposes this path as
 * C++ classes: ArmadaRTX (include/amadeus/armada.cuh:142-336), its plugin
 * interface AttachmentRTX (armada.cuh:107-139) with the PathTracer
 * (include/amadeus/path_tracer.cuh:17-167) and ReSTIR
 * (include/amadeus/restir.cuh:35-305) attachments, and the acceleration
 * structure layer amadeus::Accelerator (include/amadeus/accelerator.cuh:23-278).
 * The entry points below are the plain-pointer equivalents of those methods;
 * include/krt_armada.hpp re-creates the C++ classes on top of them.
 *
 * Conventions: every call returns KRT_OK (0) or an error code; the message of
 * the last failure on the calling thread is krt_last_error().  The reference
 * throws std::runtime_error / kobra::cuda::Exception instead
 * (include/cuda/error.cuh:43-98).  A context is single-threaded, like the
 * reference's one host thread (all launches on one stream).
 */
#ifndef KRT_H_
#define KRT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    KRT_OK = 0,
    KRT_ERR_INVALID = 1, /* bad argument or handle */
    KRT_ERR_HIP = 2,     /* HIP runtime / launch failure */
    KRT_ERR_BUDGET = 3,  /* request exceeds the HBM budget */
    KRT_ERR_STATE = 4    /* call not valid in this state (no attachment, no BVH, ...) */
};

typedef struct krt_context krt_context;
typedef struct krt_armada krt_armada;
typedef struct krt_attachment krt_attachment;
/* Scene + its acceleration structure; the analogue of OptixTraversableHandle. */
typedef uint64_t krt_handle;

/* Last error message of the calling thread ("" if none). */
const char *krt_last_error(void);

/* Context on HIP device `device`; owns one stream. */
int krt_create(int device, krt_context **out);
void krt_destroy(krt_context *ctx);

/* ---------------------------------------------------------------- scenes --
 * World-space triangle soup.  Replaces MeshDaemon::cache_cuda
 * (source/daemons/mesh.cu:25-61), the material upload of
 * ArmadaRTX::preprocess_scene (source/amadeus/armada.cu:560-572) and the
 * triangle-light list (armada.cu:98-150: every triangle whose material has
 * |emission| > 0 becomes a light, in triangle order).
 * materials: 12 floats each {diffuse.rgb, specular.rgb, emission.rgb,
 * roughness, refraction, type} with type a Shading bit set (4 = eDiffuse,
 * 2 = eTransmission; include/types.hpp:18-26). */
typedef struct {
    const float *positions; /* nv x 3 */
    const float *normals;   /* nv x 3 */
    uint32_t nv;
    const uint32_t *indices;      /* nt x 3 */
    const uint32_t *tri_material; /* nt */
    uint32_t nt;
    const float *materials; /* nm x 12 */
    uint32_t nm;
} krt_scene_desc;

int krt_scene_upload(krt_context *ctx, const krt_scene_desc *desc, krt_handle *scene);

/* Material textures and normal maps of a flat scene: calculate_material and
 * calculate_normal (source/optix/amadeus_common.cuh:74-187) with the textures
 * of convert_material (source/amadeus/armada.cu:333-398) and the vertex
 * attributes tex_coords / tangent / bitangent (include/vertex.hpp:15-20).
 * A hit's (u, v) is the interpolated tex_coords with v flipped
 * (amadeus_common.cuh:498-502); a texture is read as tex2D <float4> of an
 * RGBA8 image with normalised coordinates, bilinear filtering and wrap
 * addressing (include/cuda/interop.cuh:11-51), filtered in FP32.  Diffuse,
 * specular and emission textures replace the material's values by the
 * texel's rgb, the roughness texture by its r; a normal map tilts the shading
 * normal into the tangent frame.  Triangles of a material with an emission
 * texture are triangle lights too, with the material's constant emission
 * (armada.cu:117-150, :264-265).  The editor's ReSTIR ("mamba") reads the
 * diffuse texture only and no normal map (editor/optix/common.cuh:34-47,
 * editor/gbuffer_rtx_shader.cu:95-160). */
typedef struct {
    const float *uvs;       /* nv x 2: Vertex::tex_coords */
    const float *tangents;  /* nv x 6: Vertex::tangent, Vertex::bitangent; may be NULL without normal maps */
    uint32_t nv;            /* = the scene's nv */
    const uint32_t *texels; /* RGBA8 texels (r in the low byte), row-major, every texture back to back */
    const uint32_t *dims;   /* ntex x 2: width, height (each 1..16384) */
    uint32_t ntex;          /* < 65535 */
    const int32_t *material_textures; /* nm x 5: diffuse, normal, specular, emission, roughness texture or -1 */
    uint32_t nm;            /* = the scene's nm */
} krt_scene_textures;

int krt_scene_upload_textured(krt_context *ctx, const krt_scene_desc *desc, const krt_scene_textures *textures,
                              krt_handle *scene);

/* Two-level scene: base meshes in their own space and instances of them with
 * a row-major 3x4 transform (OptixInstance::transform).  Replaces
 * Accelerator::update (one GAS per entity submesh,
 * include/amadeus/accelerator.cuh:145-190) and build_tlas (instances,
 * :197-277).  Meshes are given back to back (indices local to their mesh).
 * Triangle ids are global: instance i's triangles follow those of instances
 * 0..i-1, in its mesh's order, and tri_material gives the material of every
 * global triangle (the reference gives each entity's submesh its own
 * material).  Hits report global ids; shading follows
 * calculate_intersection / calculate_normal with the instance transform as
 * the model matrix (source/optix/amadeus_common.cuh:62-112); lights are the
 * emissive triangles through the transform, in instance then triangle order
 * (armada.cu:117-140).  The path tracer, ReSTIR DI and the ray queries take
 * instanced scenes; ReSTIR GI, the radiance cache and the server take flat
 * scenes. */
typedef struct {
    const float *positions;     /* sum(mesh_nv) x 3 */
    const float *normals;       /* sum(mesh_nv) x 3 */
    const uint32_t *mesh_nv;    /* nmesh */
    const uint32_t *indices;    /* sum(mesh_nt) x 3, local to their mesh */
    const uint32_t *mesh_nt;    /* nmesh */
    uint32_t nmesh;
    const uint32_t *instance_mesh; /* ninst */
    const float *transforms;       /* ninst x 12, row-major 3x4, non-singular */
    uint32_t ninst;
    const uint32_t *tri_material;  /* one per global triangle */
    const float *materials;        /* nm x 12 */
    uint32_t nm;
} krt_instanced_desc;

int krt_scene_upload_instanced(krt_context *ctx, const krt_instanced_desc *desc, krt_handle *scene);
int krt_scene_free(krt_context *ctx, krt_handle scene);
int krt_scene_light_count(krt_context *ctx, krt_handle scene, uint32_t *out);
/* 12 floats per light {a, ab, ac, intensity} (include/optix/lighting.cuh:96-100) */
int krt_scene_lights(krt_context *ctx, krt_handle scene, float *out, size_t nfloats);

/* ----------------------------------------------------------- accelerator --
 * GPU build of the reference's binary "SAH" BVH (source/bvh.cpp:32-210,
 * serialised as bvh.cpp:213-260), then its collapse into the traversal layout
 * (leaves of up to leaf_max triangles, 1..32).  Replaces
 * Accelerator::build_gas/build_tlas (accelerator.cuh:46-277). */
typedef struct {
    uint32_t triangles;
    uint32_t serialized_nodes; /* 2N-1 */
    uint32_t traversal_nodes;  /* 4-wide nodes the kernels walk */
    uint32_t depth_bound;  /* upper bound of the traversal stack depth */
    uint32_t build_levels; /* level-synchronous build passes */
    float build_ms;        /* GPU time of the build (HIP events) */
    int layout;            /* KRT_LAYOUT_* of the binary tree the wide tree collapses */
    uint32_t binary_nodes; /* binary traversal tree before the 4-wide collapse */
    uint32_t node_bytes;   /* bytes fetched per traversal-node visit */
    float refit_ms;        /* GPU time of the last krt_accel_refit (0 if none) */
} krt_bvh_stats;

enum {
    KRT_LAYOUT_SAH = 0,      /* binned-SAH traversal tree (default) */
    KRT_LAYOUT_REFERENCE = 1 /* collapse of the bvh.cpp tree itself */
};
/* Builds the bvh.cpp tree (serialised layout and preorder ranks) and the
 * traversal tree of `layout`.  Either layout gives bit-identical query results:
 * every triangle carries its bvh.cpp preorder rank for the tie-break. */
int krt_accel_build_layout(krt_context *ctx, krt_handle scene, int leaf_max, int layout);
int krt_accel_build(krt_context *ctx, krt_handle scene, int leaf_max); /* layout SAH */
int krt_accel_stats(krt_context *ctx, krt_handle scene, krt_bvh_stats *out);
/* Moved geometry (Accelerator::update when an entity's transform changes,
 * accelerator.cuh:145-196; armada.cu:575-602 and :716-727 push the new
 * transform and rebuild).  krt_scene_update_vertices overwrites the world-space
 * positions (and normals, when not NULL) of vertices [first, first + count)
 * of a flat scene; the acceleration structure is then out of date and every
 * render or query of the scene returns KRT_ERR_STATE until krt_accel_refit
 * (same tree shapes, boxes recomputed bottom-up from the new vertices, the
 * triangle records and the triangle lights rewritten) or krt_accel_build.
 * Either bumps the scene's version, so an armada restarts accumulation
 * (armada.cu:771-772).  Both refuse a scene held by server requests. */
int krt_scene_update_vertices(krt_context *ctx, krt_handle scene, uint32_t first, uint32_t count,
                              const float *positions, const float *normals);
int krt_accel_refit(krt_context *ctx, krt_handle scene);
/* bvh.cpp serialised layout, 12 floats (3 x vec4) per node */
int krt_accel_serialized(krt_context *ctx, krt_handle scene, float *out, size_t nfloats);
/* primitive id of each leaf in preorder (the closest-hit tie-break rank) */
int krt_accel_leaf_order(krt_context *ctx, krt_handle scene, int32_t *out, size_t n);
/* Instanced scenes: the bvh.cpp tree of one base mesh, and the bvh.cpp tree
 * over the instances' world boxes (the 8 corners of the mesh's root box
 * through the transform, widened by 1e-4 of the largest extent + 1e-5 of the
 * largest coordinate); 12 floats per node, *nodes = the node count (out may
 * be NULL to query it).  Instanced closest hits are the lexicographic
 * minimum of (t, instance rank in this tree, triangle rank in its mesh's). */
int krt_accel_mesh_serialized(krt_context *ctx, krt_handle scene, uint32_t mesh, float *out, size_t nfloats,
                              uint32_t *nodes);
int krt_accel_instance_tree(krt_context *ctx, krt_handle scene, float *out, size_t nfloats, uint32_t *nodes);

/* Ray queries on host buffers.  rays: 8 floats {o.xyz, tmax, d.xyz, pad};
 * hits: 4 words {t, b1, b2, prim} with prim = -1 on a miss.  The closest hit
 * is the lexicographic minimum of (t, preorder leaf rank) over triangles with
 * 0 < t < tmax (Möller–Trumbore, shaders/rt/modules/intersect.glsl:14-33).
 * Replaces optixTrace (source/optix/amadeus_common.cuh:467-478, :190-212). */
int krt_trace_closest(krt_context *ctx, krt_handle scene, const float *rays, uint32_t n, float *hits);
int krt_trace_occluded(krt_context *ctx, krt_handle scene, const float *rays, uint32_t n, int32_t *occluded);
/* Same on device buffers (rays/hits already in HBM); enqueued on the
 * context stream, no synchronisation. */
int krt_trace_closest_device(krt_context *ctx, krt_handle scene, const void *rays_dev, uint32_t n, void *hits_dev);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same query
 * and the same hits as krt_trace_closest_device, but the traversal stack's
 * overflow slab is a fixed 16 MiB device array of the library (one per GPU,
 * held until the process ends) instead of memory allocated on demand, so
 * the call makes no hipMalloc / hipFree.  The workgroup count is cut to what
 * the slab holds (tree depth bound x 128 lanes per workgroup);
 * KRT_ERR_BUDGET if the tree is deeper than the slab allows.  Calls that
 * overlap on one GPU (several contexts or threads) share the slab and race. */
int krt_trace_closest_device_noalloc(krt_context *ctx, krt_handle scene, const void *rays_dev, uint32_t n,
                                     void *hits_dev);

/* Streaming copy of `bytes` between two device buffers, `reps` times: average
 * milliseconds per copy (HBM bandwidth and profiler byte-counter calibration;
 * no reference counterpart). */
int krt_copy_probe(krt_context *ctx, size_t bytes, int reps, float *ms);

/* HIP events on the context's stream around the work queued between the two
 * calls: milliseconds of device time (no reference counterpart). */
int krt_timer_begin(krt_context *ctx);
int krt_timer_end(krt_context *ctx, float *ms);
/* Waits for all work queued on the context's stream. */
int krt_sync(krt_context *ctx);

/* Device buffers for the *_device entry points (no reference counterpart):
 * allocation, and stream-ordered copies that return when complete. */
int krt_buffer_create(krt_context *ctx, size_t bytes, void **dev);
int krt_buffer_destroy(krt_context *ctx, void *dev);
int krt_buffer_upload(krt_context *ctx, void *dev, const void *host, size_t bytes);
int krt_buffer_download(krt_context *ctx, void *host, const void *dev, size_t bytes);

/* ---------------------------------------------------------------- render --
 * ArmadaRTX (armada.cuh:142-336) and its attachments. */
typedef struct {
    float center[3]; /* ArmadaLaunchInfo::Camera (armada.cuh:46-55) */
    float U[3], V[3], W[3];
    float pv[16]; /* projection * view, column-major (for reprojection) */
} krt_camera;

/* Camera rays of make_ray (source/optix/amadeus_path_tracer.cu:26-56, jittered;
 * amadeus_restir.cu:31-65, fixed -0.5 offset when jitter == 0) for frame
 * `sample`, 8 floats per pixel as krt_trace_closest takes them. */
int krt_camera_rays(krt_context *ctx, const krt_camera *cam, uint32_t width, uint32_t height, int jitter,
                    int sample, uint32_t user_seed, float *rays);

enum { KRT_OPT_NONE = 0, KRT_OPT_INT = 1, KRT_OPT_FLOAT = 2, KRT_OPT_BOOL = 3, KRT_OPT_STRING = 4 };
/* AttachmentRTX::OptionValue = variant<int, float, bool, string> (armada.cuh:120) */
typedef struct {
    int kind;
    int i;
    float f;
    int b;
    const char *s;
} krt_option;

typedef struct {
    void *color, *normal, *albedo, *position; /* W x H float4 each (armada.cu:70-73) */
} krt_buffers;

typedef struct {
    uint64_t closest_rays; /* camera + bounce rays traced */
    uint64_t shadow_rays;  /* visibility rays traced */
    uint64_t samples;      /* path samples (pixels x frames) */
    float gpu_ms;          /* GPU time of the last render call (HIP events) */
    float kernel_ms[8];    /* per-pass GPU time: [0] path tracer / initial, [1] temporal, [2] spatial */
    uint64_t pass_rays[4]; /* rays traced per pass, same indexing */
    uint64_t node_visits;  /* traversal nodes fetched by all rays of the frame */
    uint64_t tri_tests;    /* ray-triangle tests of the frame */
    uint64_t pass_nodes[4];
    uint64_t pass_tris[4];
} krt_frame_stats;

/* ArmadaRTX(context, accelerator, mesh daemon, extent) (armada.cu:28-83) */
int krt_armada_create(krt_context *ctx, uint32_t width, uint32_t height, krt_armada **out);
void krt_armada_destroy(krt_armada *a);
/* Attachment factory: kind "path_tracer" | "restir" | "restir_gi" |
 * "mamba" (the editor's G-buffer ReSTIR DI, editor/optix/mamba_shader.cu).
 * Lifetimes: destroy armadas before the attachments attached to them, and
 * everything created on a context before krt_destroy. */
int krt_attachment_create(krt_context *ctx, const char *kind, krt_attachment **out);
void krt_attachment_destroy(krt_attachment *att);
/* ArmadaRTX::attach / activate / set_option / get_option / set_depth
 * (armada.cuh:275-320) */
int krt_armada_attach(krt_armada *a, const char *name, krt_attachment *att);
int krt_armada_activate(krt_armada *a, const char *name);
int krt_armada_set_option(krt_armada *a, const char *field, const krt_option *value);
int krt_armada_get_option(krt_armada *a, const char *field, krt_option *out);
int krt_armada_set_depth(krt_armada *a, int depth);
/* ArmadaRTX::set_envmap / set_envmap_enabled (armada.cu:86-96,
 * armada.cuh:323-328): the environment map as width x height RGBA float
 * texels, row-major (row 0 is v = 0), copied to the device; rgba == NULL or a
 * zero extent removes it.  Sampled like the reference's texture (normalised
 * coordinates, bilinear, wrap): rays that miss take its radiance (the path
 * tracer at (u, 1 - v), ReSTIR at (u, v)), and next-event estimation and
 * ReSTIR's candidates pick it as one more light.  restir_gi refuses a frame
 * while it is enabled (KRT_ERR_STATE); the editor's mamba attachment does not
 * read it.  Setting it does not restart accumulation, as in the reference. */
int krt_armada_set_envmap(krt_armada *a, const float *rgba, uint32_t width, uint32_t height);
int krt_armada_set_envmap_enabled(krt_armada *a, int enabled);
/* ArmadaRTX::render(system, entities, transforms, camera, transform,
 * accumulate) (armada.cu:733-780): one frame of the active attachment. */
int krt_armada_render(krt_armada *a, krt_handle scene, const krt_camera *camera, int accumulate);
/* color/normal/albedo/position_buffer() (armada.cuh:287-301) */
int krt_armada_buffers(krt_armada *a, krt_buffers *out);
/* which: 0 color, 1 normal, 2 albedo, 3 position; out: W*H*4 floats */
int krt_armada_readback(krt_armada *a, int which, float *out);
int krt_armada_upload(krt_armada *a, int which, const float *in);
int krt_armada_stats(krt_armada *a, krt_frame_stats *out);
/* Converged-reference accumulator (no reference counterpart; the reference
 * accumulates only the FP32 running mean of accumulate(), amadeus_path_tracer.cu:
 * 59-70): while enabled, every path-tracer sample and its square are also
 * added to FP64 sums per pixel.  Enabling (again) clears the sums.  read: 3
 * doubles per pixel and the number of frames summed; read_f64_squares: the 3
 * sums of squares per pixel (the per-pixel sample variance of one frame is
 * squares/n - (sums/n)^2). */
int krt_armada_set_f64_accumulation(krt_armada *a, int enable);
/* ReSTIR reservoirs (include/amadeus/reservoir.cuh:19-83) as 16 floats per
 * pixel {Le.xyz, normal.xyz, point.xyz, type, w, W, M, seed.xyz}; for
 * "restir_gi" 20 floats {x_s, n_s, L_o, x_owner, w, W, M, seed.xyz, 0, 0}.
 * which 0: the buffer the next frame's temporal pass merges with (last
 * spatial output), 1: the other one.  For "mamba" 16 floats {light position,
 * normal, emission, area, w, W, M, seed}; which 0: the spatial output of the
 * last frame, 1: the temporal output the next frame reads. */
int krt_armada_read_reservoirs(krt_armada *a, int which, float *out);
int krt_armada_read_f64(krt_armada *a, double *out, uint64_t *frames);
int krt_armada_read_f64_squares(krt_armada *a, double *out);
int krt_armada_sync(krt_armada *a);

/* ---------------------------------------------------------- tonemapping --
 * hdr_to_ldr (include/cuda/color.cuh:82-121, called by bench.cu:205 and
 * motion_capture.cuh:358 on the colour buffer): float4 HDR pixels ->
 * RGBA8 packed in a uint32 as r | g << 8 | b << 16 | 255 << 24
 * (include/cuda/cast.cuh:16-20), rows flipped vertically (color.cuh:92-93),
 * through ACES (color.cuh:35-41) or sRGB (color.cuh:14-31), each channel
 * quantized by quantize_unsigned_8_bits (color.cuh:43-56).  Device pointers
 * (width x height float4 in, width x height uint32 out); enqueued on the
 * context stream without synchronisation, as the reference's launch. */
enum { KRT_TONEMAP_ACES = 0, KRT_TONEMAP_SRGB = 1 };
int krt_hdr_to_ldr(krt_context *ctx, const void *pixels_dev, void *target_dev, uint32_t width, uint32_t height,
                   int tonemapping);
/* The armada's colour buffer through krt_hdr_to_ldr into a host array of
 * W*H uint32; returns when the copy is complete. */
int krt_armada_read_ldr(krt_armada *a, int tonemapping, uint32_t *out);
/* Image output (the still-image capture of bench.cu:226-233, stbi_write_png
 * of the hdr_to_ldr frame: RGBA8, rows in memory order): a PNG file of
 * width x height packed pixels, and of the armada's colour buffer through
 * krt_hdr_to_ldr.  Video capture (include/capture.hpp, ffmpeg) is not
 * provided. */
int krt_write_png(const char *path, const uint32_t *rgba, uint32_t width, uint32_t height);
/* Image input for material textures (the loader behind
 * TextureLoader::load_texture, source/image.cpp:96-118, which takes stbi_load
 * with 4 channels): a PNG file of any colour type and bit depth (16-bit
 * samples keep their high byte), Adam7 interlacing and tRNS transparency, as
 * packed RGBA8 (r in the low byte), rows in file order.  With rgba == NULL
 * only width and height are read; capacity counts uint32 texels. */
int krt_read_png(const char *path, uint32_t *width, uint32_t *height, uint32_t *rgba, size_t capacity);
int krt_armada_write_png(krt_armada *a, int tonemapping, const char *path);

/* Plugin entry of source/amadeus/naive_path_tracer.cu:280-344
 * (`extern "C" ret load_attachment()`), by kind. */
typedef struct {
    const char *name;
    krt_attachment *ptr;
} krt_attachment_ret;
krt_attachment_ret krt_load_attachment(krt_context *ctx, const char *kind);

/* ------------------------------------------------------- radiance cache --
 * The online-trained radiance-cache MLP.  It takes the place of
 * editor/mlp.cuh:71-154 (RadianceField, dev_radiance_field, forward<B>,
 * release), whose network is a never-compiled FP32 forward pass; here
 * 62 encoded inputs -> 5 x 64 ReLU -> 3 outputs, BF16 on MFMA with FP32
 * accumulation, trained by Adam with dynamic loss scaling.
 * Queries: KRT_MLP_RAW floats each {pos.xyz, dir.xyz, normal.xyz,
 * albedo.rgb, roughness}; pos is mapped to [0,1] by
 * (pos - pos_min) * pos_scale before encoding.  Outputs / targets: 3 floats
 * (RGB radiance).  Parameters: KRT_MLP_PARAMS floats, per layer W[out][in]
 * row-major then b[out] (mlp.cuh:48-67 ordering). */
#define KRT_MLP_RAW 13
#define KRT_MLP_OUT 3
#define KRT_MLP_PARAMS 20867

typedef struct krt_mlp krt_mlp;

typedef struct {
    float lr, beta1, beta2, eps;  /* Adam */
    float loss_scale;             /* initial loss scale */
    int32_t growth_interval;      /* finite steps before the scale doubles */
    float pos_min[3], pos_scale[3];
} krt_mlp_config;

typedef struct {
    float loss_scale, beta1_t, beta2_t, last_loss;
    int32_t step, good_steps, skipped_steps, pending_nonfinite;
} krt_mlp_state;

void krt_mlp_default_config(krt_mlp_config *cfg);
/* dev_radiance_field (mlp.cuh:109-130): weights and biases
 * xavier(n, m) * (u - 0.5) with u from a seeded SplitMix64 stream. */
int krt_mlp_create(krt_context *ctx, const krt_mlp_config *cfg, uint64_t seed, krt_mlp **out);
void krt_mlp_destroy(krt_mlp *net); /* release (mlp.cuh:132-150) */
int krt_mlp_get_params(krt_mlp *net, float *params);
int krt_mlp_set_params(krt_mlp *net, const float *params);
int krt_mlp_set_encoding(krt_mlp *net, const float pos_min[3], const float pos_scale[3]);
/* forward<B> (mlp.cuh:153-163) for n queries; host or device arrays. */
int krt_mlp_infer(krt_mlp *net, const float *raw, uint32_t n, float *out);
int krt_mlp_infer_device(krt_mlp *net, const float *raw_dev, uint32_t n, float *out_dev);
/* One optimiser step on n (query, target) pairs: relative L2 loss
 * (y - t)^2 / (y^2 + 0.01) averaged over 3n outputs.  The host variant
 * returns the batch loss; the device variant queues the step and returns. */
int krt_mlp_train_step(krt_mlp *net, const float *raw, const float *target, uint32_t n, float *loss);
int krt_mlp_train_step_device(krt_mlp *net, const float *raw_dev, const float *target_dev, uint32_t n);
/* Gradient of the batch loss (unscaled, canonical order) without a step, and
 * an Adam step with a caller-supplied gradient (e.g. after a cross-device
 * reduction). */
int krt_mlp_gradient(krt_mlp *net, const float *raw, const float *target, uint32_t n, float *grad,
                     float *loss);
int krt_mlp_apply_gradient(krt_mlp *net, const float *grad);
int krt_mlp_get_state(krt_mlp *net, krt_mlp_state *state);
/* The network of a path_tracer attachment's radiance cache (options
 * "radiance_cache", "cache_depth", "train_samples", "cache_training"):
 * a borrowed handle, NULL before the first cached frame, valid until the
 * attachment is unloaded or destroyed; krt_mlp_destroy ignores it. */
int krt_attachment_cache(krt_attachment *att, krt_mlp **out);

/* ------------------------------------------------------------- serving --
 * Render-request server (no reference counterpart): continuous batching of
 * path-tracer and ReSTIR DI requests.  A request's answer is the colour
 * buffer a path_tracer (or restir) attachment holds after `spp` accumulated
 * frames of the same scene, camera, depth, seed and options
 * (ArmadaRTX::render, armada.cu:733-780), bit for bit, whatever else runs
 * beside it.  Every step advances each active request by its share of
 * `step_pixels` pixel-frames (at least one frame): small path-tracer
 * requests in one persistent launch (a captured hipGraph); large ones
 * (`wave_pixels`) through the wavefront passes, and each ReSTIR request's
 * frames, on one of a few streams so that they overlap on the device, each
 * such request's reservoirs or wavefront buffers resident from admission to
 * its last frame.  Requests are
 * admitted FIFO as slots and HBM budget free up.  While a request holds a
 * scene, krt_scene_free and krt_accel_build of that scene fail with
 * KRT_ERR_STATE. */
typedef struct krt_server krt_server;

typedef struct {
    uint64_t hbm_budget;   /* bytes: fixed scratch + each active request's buffers (64 B per pixel for a
                              batched path-tracer request, plus 160 B + 49 B per path vertex for one on a
                              lane, about 0.9 KB per pixel for ReSTIR) */
    uint32_t max_active;   /* requests per step, 1..64 */
    int32_t use_graph;     /* 1: replay a captured hipGraph per step */
    double max_queue_ms;   /* > 0: reject requests queued longer than this */
    uint64_t step_pixels;  /* pixel-frames per step, shared fairly among active requests */
    uint64_t wave_pixels;  /* path-tracer requests of at least this many pixels run the wavefront passes on a
                              lane, smaller ones the persistent batched launch (0: every request on lanes) */
} krt_server_config;

enum { KRT_REQUEST_PATH_TRACER = 0, KRT_REQUEST_RESTIR = 1 };

/* The restir attachment's options ("candidates", "m_cap", "spatial_passes",
 * "neighbours", "radius"; include/amadeus/restir.cuh:223-304 runs 32, 200, 2,
 * 5, 30). */
typedef struct {
    int32_t candidates, m_cap, spatial_passes, neighbours, radius;
} krt_restir_options;

typedef struct {
    krt_handle scene;
    krt_camera camera;
    uint32_t width, height, spp; /* spp: accumulated frames */
    int32_t max_depth;     /* 0..8 */
    uint32_t seed;         /* the attachment's "seed" option */
    int32_t kind;          /* KRT_REQUEST_PATH_TRACER or KRT_REQUEST_RESTIR */
    krt_restir_options restir; /* KRT_REQUEST_RESTIR only */
} krt_request;

typedef struct {
    uint64_t submitted, completed, rejected, steps;
    uint32_t active, queued, max_concurrent, reserved;
    double p50_ms, p99_ms;           /* request latency, submit -> result ready */
    double step_p50_ms, step_p99_ms; /* device time of one batched step */
    double samples;                  /* pixels x spp of completed requests */
    double busy_ms, wall_ms;         /* device time of all steps; first submit -> last completion */
    uint64_t hbm_scratch, hbm_in_use, hbm_peak;
} krt_server_stats;

void krt_server_default_config(krt_server_config *cfg);
int krt_server_create(krt_context *ctx, const krt_server_config *cfg, krt_server **out);
void krt_server_destroy(krt_server *srv);
/* KRT_ERR_BUDGET when the request alone exceeds the budget. */
int krt_server_submit(krt_server *srv, const krt_request *req, uint64_t *id);
/* Blocks until the request completes (color: width x height x 4 floats, may be
 * NULL) or was rejected (KRT_ERR_BUDGET); the id is released afterwards. */
int krt_server_wait(krt_server *srv, uint64_t id, float *color, double *latency_ms);
int krt_server_drain(krt_server *srv);
int krt_server_get_stats(krt_server *srv, krt_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif

/*
 * CPU oracle for the ray-traced rendering path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17 restatement of the reference renderer's algorithms:
 *   - per-triangle AABBs           include/mesh.hpp:48-78
 *   - "SAH" binary BVH build       source/bvh.cpp:32-210
 *   - threaded serialisation       source/bvh.cpp:213-260
 *   - stackless traversal          shaders/rt/modules/bvh.glsl:52-127,
 *                                  shaders/rt/modules/bbox.glsl:8-66
 *   - Möller–Trumbore              shaders/rt/modules/intersect.glsl:14-33
 *   - pcg RNG, GGX/Glass BSDF      include/cuda/random.cuh, include/cuda/brdf.cuh
 *   - NEE with triangle lights     source/optix/amadeus_common.cuh:190-464,
 *                                  include/optix/lighting.cuh:96-174
 *   - path tracer estimator        source/optix/amadeus_path_tracer.cu:26-228
 *   - ReSTIR DI passes             source/optix/amadeus_restir.cu:82-559,
 *                                  include/amadeus/reservoir.cuh:19-76,
 *                                  include/amadeus/restir.cuh:223-304
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so, and only as the checker.  The product never links it.
 *
 * Pinning (what each function is checked against, tests/test_oracle_*.py):
 *   pcg3d/pcg3f      JCGT 2020 integer hash recomputed in numpy + committed
 *                    known-answer vectors.
 *   bvh build        hand-computed trees for N = 1, 2, 3, 5; SAH cost closed
 *                    form; invariants (2N-1 nodes, containment, threading).
 *   traversal        brute force over all triangles, adversarial rays.
 *   BSDF / lights    FP64 quadrature of the pdf, light-sample moments,
 *                    closed-form irradiance of one small triangle light.
 *   PT estimator     furnace-style closed forms on one-light scenes + the
 *                    FP64 relMSE of a converged image (GPU tests).
 *   ReSTIR DI        RIS unbiasedness with one light (closed form); reservoir
 *                    stream known answers.
 */
#ifndef KRT_ORACLE_H_
#define KRT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene: flattened world-space triangle soup.  mats: 12 floats per material
 * {diffuse.rgb, specular.rgb, emission.rgb, roughness, refraction, type}. */
typedef struct orc_scene orc_scene;

orc_scene *orc_scene_create(const float *pos, const float *nrm, int nv,
                            const uint32_t *idx, const uint32_t *tri_mat, int nt,
                            const float *mats, int nm,
                            const float *bvh_nodes, int nnodes);
/* Two-level scene (Accelerator::update / build_tlas,
 * include/amadeus/accelerator.cuh:145-277): nmesh base meshes given as
 * concatenated arrays (indices local to their mesh), ninst instances
 * {mesh, row-major 3x4 transform}.  Builds every mesh's bvh.cpp tree and the
 * bvh.cpp tree over the instances' padded world boxes (orc_instance_box).
 * Triangle ids are global: instance i's triangles follow those of instances
 * 0..i-1; tri_mat gives the material of every global triangle. */
orc_scene *orc_scene_create_instanced(int nmesh, const int *mesh_nv, const int *mesh_nt, const float *pos,
                                      const float *nrm, const uint32_t *idx, int ninst, const int *inst_mesh,
                                      const float *inst_xform, const uint32_t *tri_mat, const float *mats, int nm);
/* the instance tree (12 floats per node, returns the node count) */
int orc_scene_tlas(const orc_scene *s, float *nodes);
void orc_affine_inverse(const float *M, float *Minv);
void orc_instance_box(const float *root_min, const float *root_max, const float *M, float *box6);
void orc_scene_destroy(orc_scene *s);
/* Material textures of a flat scene (calculate_material / calculate_normal,
 * amadeus_common.cuh:74-187): nv x 2 tex_coords, nv x 6 {tangent, bitangent}
 * (may be NULL), RGBA8 texels of ntex textures back to back (dims: w, h
 * each), 5 texture ids per material {diffuse, normal, specular, emission,
 * roughness} (-1: none).  Rebuilds the triangle lights. */
void orc_scene_set_textures(orc_scene *s, const float *uvs, const float *tangents, const uint32_t *texels,
                            const uint32_t *dims, int ntex, const int32_t *mat_tex);
/* The texture fetch, for its pins: n lookups (u, v) of one w x h RGBA8
 * texture, 4 floats each. */
void orc_tex_lookup(const uint32_t *texels, uint32_t w, uint32_t h, const float *uv, int n, float *out);
/* The armada's environment map (ArmadaRTX::set_envmap): w x h RGBA float
 * texels, row-major; rgba == NULL removes it. */
void orc_scene_set_envmap(orc_scene *s, const float *rgba, uint32_t w, uint32_t h);
/* The environment map's math, for its pins: which 0 = atan2(in[2i], in[2i+1]),
 * 1 = asin(in[2i]); the radiance along n directions (3 floats each); n
 * directions of Ld_Environment's sampling from one seed. */
void orc_env_math(const float *in, int n, int which, float *out);
void orc_env_lookup(const orc_scene *s, const float *dirs, int n, int flip_v, float *out);
void orc_env_sample_dirs(uint32_t seed_word, int n, float *out);
int orc_scene_light_count(const orc_scene *s);
/* 12 floats per light: a, ab, ac, intensity (armada.cu:98-150 rule). */
void orc_scene_lights(const orc_scene *s, float *out);

void orc_pcg3d(const uint32_t *in, uint32_t *out, int n);
/* n seeds of 3 floats, advanced in place by one pcg3f step each. */
void orc_pcg3f(float *seeds, int n);
void orc_sincos_turn(const float *u, float *c, float *s, int n);

void orc_tri_boxes(const float *pos, const uint32_t *idx, int nt, float *boxes);
/* Returns the node count (2n-1), writes 12 floats per node. */
int orc_bvh_build(const float *boxes, int n, float *nodes);
/* Leaf objects in preorder (the tie-break rank of a primitive). */
int orc_bvh_leaf_order(const float *nodes, int nnodes, int32_t *order);
int orc_bvh_depth(const float *nodes, int nnodes);

/* rays: 8 floats {o.xyz, tmax, d.xyz, pad}; hits: 4 words {t, b1, b2, prim}
 * mode 0 brute force, 1 GLSL stackless walk (exact reference box test),
 * 2 stackless walk with a conservative box test.
 * counters[0] += node visits, counters[1] += triangle tests. */
void orc_trace(const orc_scene *s, const float *rays, int n, int mode, float *hits,
               uint64_t *counters);
void orc_occluded(const orc_scene *s, const float *rays, int n, int32_t *out);

/* BSDF probes for tests: sh = 12 floats {n, wo, diffuse, specular}, mat
 * params roughness; returns pdf / brdf / sample. */
float orc_ggx_pdf(const float *n, const float *wi, const float *wo, const float *kd,
                  const float *ks, float roughness);
void orc_ggx_brdf(const float *n, const float *wi, const float *wo, const float *kd,
                  const float *ks, float roughness, float *out);
void orc_ggx_sample(const float *n, const float *wo, const float *kd, const float *ks,
                    float roughness, float *seed, float *wi_out);
void orc_light_sample(const float *light12, float *seed, float *point_out);
/* count GGX/Lambert samples (brdf.cuh:386-433) from the seeds (i, word, 0x51) */
void orc_ggx_sample_batch(const float *n, const float *wo, const float *kd, const float *ks, float roughness,
                          int count, uint32_t word, float *wi_out);
/* Glass::sample direction (brdf.cuh:144-175) for incident direction wi */
void orc_glass_sample(const float *n, const float *wi, float eta, int entering, float *out);

typedef struct {
    float center[3], U[3], V[3], W[3];
    float pv[16]; /* projection*view of this camera, column-major */
} orc_camera;

typedef struct {
    int width, height;
    int max_depth;
    int samples;    /* accumulated-frame counter (ArmadaLaunchInfo::samples) */
    int accumulate;
    uint32_t user_seed;
    /* ReSTIR */
    int candidates, m_cap, spatial_passes, neighbours, radius;
    int fixed_offset; /* path tracer camera rays at the fixed -0.5 offset (as the ReSTIR raygen) */
} orc_params;

/* counters: [0] closest rays, [1] shadow rays, [2] discarded depth-cut rays,
 * [3] node visits, [4] triangle tests */
void orc_render_pt(const orc_scene *s, const orc_camera *cam, const orc_params *p,
                   float *color, float *normal, float *albedo, float *position,
                   int nthreads, uint64_t *counters);

/* Same as orc_render_pt for a list of pixels only (one sample each): used to
 * spot-check large frames.  out: 4 floats per pixel (the un-accumulated
 * color sample). */
void orc_render_pt_pixels(const orc_scene *s, const orc_camera *cam, const orc_params *p,
                          const int32_t *pix, int n, float *out);

/* Closed-form-friendly FP64 accumulation of many frames into acc (3 doubles
 * per pixel): renders frames [first, first+count) and adds each sample. */
void orc_accumulate_pt_f64(const orc_scene *s, const orc_camera *cam, const orc_params *p,
                           int first, int count, double *acc, int nthreads);

typedef struct orc_restir orc_restir;
orc_restir *orc_restir_create(int width, int height, uint32_t user_seed);
void orc_restir_destroy(orc_restir *r);
void orc_restir_frame(orc_restir *r, const orc_scene *s, const orc_camera *cam,
                      const orc_params *p, float *color, float *normal, float *albedo,
                      float *position, int nthreads, uint64_t *counters);
/* 16 floats per pixel: Le.xyz, normal.xyz, point.xyz, type, w, W, M, seed.xyz
 * which: 0 = reservoir the next frame merges with (the last spatial output),
 *        1 = the other buffer */
void orc_restir_reservoirs(const orc_restir *r, int which, float *out);

/* ReSTIR GI (our algorithm; see oracle.cpp).  color: float4 per pixel,
 * accumulated like the other attachments.  Reservoirs: 20 floats per pixel
 * {xs, ns, Lo, xo, w, W, M, seed, 0, 0}; which as orc_restir_reservoirs. */
typedef struct orc_gi orc_gi;
orc_gi *orc_gi_create(int width, int height, uint32_t user_seed);
void orc_gi_destroy(orc_gi *r);
void orc_gi_frame(orc_gi *r, const orc_scene *s, const orc_camera *cam, const orc_params *p, float *color,
                  int nthreads, uint64_t *counters);
void orc_gi_reservoirs(const orc_gi *r, int which, float *out);
/* camera rays of make_ray (8 floats each), jittered (path tracer) or fixed -0.5 offset (ReSTIR) */
void orc_camera_rays(const orc_camera *cam, const orc_params *p, int jitter, float *rays);

/* Radiance-cache MLP (mlp.cpp): 13 raw inputs per query {pos, dir, normal,
 * albedo, roughness}, pos mapped to [0,1] by (pos - pmin) * pscale; 64
 * features (62 used); params in canonical order, per layer W[out][in] then
 * b[out].  bf16 != 0 selects the kernels' rounding points. */
int orc_mlp_param_count(void);
void orc_mlp_encode(const float *raw, int n, const float *pmin, const float *pscale, float *feat);
void orc_mlp_forward(const double *params, const float *raw, int n, const float *pmin, const float *pscale,
                     int bf16, double *out);
double orc_mlp_loss(const double *params, const float *raw, const float *target, int n, const float *pmin,
                    const float *pscale, const double *den);
double orc_mlp_grad(const double *params, const float *raw, const float *target, int n, const float *pmin,
                    const float *pscale, int bf16, float loss_scale, double *grad);
void orc_mlp_adam(float *params, float *m, float *v, const float *grad, int n, float lr, float b1, float b2,
                  float eps, float *b1t, float *b2t);
/* reconnection-shift Jacobian d(omega_r)/d(omega_q) (0 when rejected) */
float orc_gi_jacobian(const float *xq, const float *xr, const float *xs, const float *ns);

/* Mamba, the editor's G-buffer ReSTIR DI (editor/optix/mamba_shader.cu): one
 * frame = ray-cast G-buffer, initial pass (8 candidates), temporal reuse if
 * `temporal`, spatial reuse (2 neighbours, 20 px) if `spatial`; color:
 * float4 per pixel, the accumulated direct lighting.  Reservoirs: 16 floats
 * per pixel {position, normal, emission, area, w, W, M, seed}; which 0 = the
 * current reservoirs, 1 = the previous frame's. */
typedef struct orc_mamba orc_mamba;
orc_mamba *orc_mamba_create(const orc_scene *s, int width, int height, uint32_t user_seed);
void orc_mamba_destroy(orc_mamba *m);
int orc_mamba_light_count(const orc_mamba *m);
void orc_mamba_frame(orc_mamba *m, const orc_scene *s, const orc_camera *cam, int temporal, int spatial,
                     int accumulate, float *color, int nthreads);
void orc_mamba_reservoirs(const orc_mamba *m, int which, float *out);

/* hdr_to_ldr (include/cuda/color.cuh:82-121): float4 pixels -> packed RGBA8,
 * rows flipped; tonemapping 0 = ACES (:35-41), 1 = sRGB (:14-31). */
void orc_hdr_to_ldr(const float *pixels, uint32_t *target, int width, int height, int tonemapping);

#ifdef __cplusplus
}
#endif

#endif

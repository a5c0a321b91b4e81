// CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <map>
#include <thread>
#include <vector>

#include "rt.hpp"

using namespace orc;

namespace {

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
struct Hit {
    float t, b1, b2;
    int prim;
};

struct Counters {
    uint64_t closest = 0, shadow = 0, discarded = 0, nodes = 0, tris = 0;
    void add(const Counters &o) {
        closest += o.closest;
        shadow += o.shadow;
        discarded += o.discarded;
        nodes += o.nodes;
        tris += o.tris;
    }
};

}  // namespace

// One mesh in its own space with its bvh.cpp tree (a flat scene is one such
// mesh in world space; an instanced scene has one per base mesh).
struct OrcMesh {
    std::vector<f3> pos, nrm;
    std::vector<uint32_t> idx;
    std::vector<float> nodes;  // 12 floats per node (3 x vec4)
    int nnodes = 0;
    std::vector<int32_t> leaf_order;
};

// Instance of a base mesh: row-major 3x4 transform (OptixInstance::transform,
// include/amadeus/accelerator.cuh:204-212), its inverse, the global id of the
// mesh's first triangle and its rank in the preorder of the instance tree.
// Materials are per global triangle (orc_scene::tri_mat), as the reference
// gives every entity's submesh its own material and its own GAS.
struct OrcInst {
    int mesh;
    float M[12], Minv[12];
    uint32_t prim_base;
    int rank;
};

struct orc_scene {
    std::vector<f3> pos, nrm;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<Material> mats;
    std::vector<float> nodes;  // 12 floats per node (3 x vec4)
    int nnodes = 0;
    std::vector<int32_t> leaf_order;
    std::vector<TriLight> lights;
    // two-level scenes (Accelerator::update / build_tlas, accelerator.cuh:145-277)
    bool instanced = false;
    std::vector<OrcMesh> meshes;
    std::vector<OrcInst> inst;
    std::vector<float> tlas;  // bvh.cpp tree over the instances' world boxes
    int ntlas = 0;
    std::vector<int32_t> inst_order;  // instances in preorder of the instance tree
    // material textures (orc_scene_set_textures): per-vertex tex_coords and
    // {tangent, bitangent}, RGBA8 texels and {offset, w, h} per texture, 5
    // texture ids per material {diffuse, normal, specular, emission,
    // roughness} (-1: none)
    bool textured = false;
    std::vector<float> uv, tan;
    std::vector<uint32_t> texels, tex_desc;
    std::vector<int32_t> mat_tex;
    Texture texture(int32_t id) const {
        Texture t;
        t.texels = texels.data() + tex_desc[3 * id];
        t.w = tex_desc[3 * id + 1];
        t.h = tex_desc[3 * id + 2];
        return t;
    }
    // environment map (the armada's, ArmadaRTX::set_envmap): RGBA texels
    std::vector<float> env;
    uint32_t env_w = 0, env_h = 0;
    EnvMap envmap() const {
        EnvMap e;
        if (env_w > 0) {
            e.tex = env.data();
            e.w = env_w;
            e.h = env_h;
        }
        return e;
    }
};

static inline f3 load3(const float *p) { return mk(p[0], p[1], p[2]); }

// glm mat4 x vec4 (column sums paired as (c0 x + c1 y) + (c2 z + c3 w)) for a
// row-major 3x4 affine transform; w = 1 for points, 0 for directions.
static inline f3 xform_point(const float *M, f3 p) {
    return mk((M[0] * p.x + M[1] * p.y) + (M[2] * p.z + M[3] * 1.0f),
              (M[4] * p.x + M[5] * p.y) + (M[6] * p.z + M[7] * 1.0f),
              (M[8] * p.x + M[9] * p.y) + (M[10] * p.z + M[11] * 1.0f));
}
static inline f3 xform_dir(const float *M, f3 v) {
    return mk((M[0] * v.x + M[1] * v.y) + (M[2] * v.z + M[3] * 0.0f),
              (M[4] * v.x + M[5] * v.y) + (M[6] * v.z + M[7] * 0.0f),
              (M[8] * v.x + M[9] * v.y) + (M[10] * v.z + M[11] * 0.0f));
}



orc_scene *orc_scene_create(const float *pos, const float *nrm, int nv, const uint32_t *idx,
                            const uint32_t *tri_mat, int nt, const float *mats, int nm,
                            const float *bvh_nodes, int nnodes) {
    orc_scene *s = new orc_scene();
    s->pos.resize(nv);
    s->nrm.resize(nv);
    for (int i = 0; i < nv; i++) {
        s->pos[i] = load3(pos + 3 * i);
        s->nrm[i] = load3(nrm + 3 * i);
    }
    s->idx.assign(idx, idx + 3 * (size_t)nt);
    s->tri_mat.assign(tri_mat, tri_mat + nt);
    s->mats.resize(nm);
    for (int m = 0; m < nm; m++) {
        const float *q = mats + 12 * m;
        Material &M = s->mats[m];
        M.diffuse = load3(q);
        M.specular = load3(q + 3);
        M.emission = load3(q + 6);
        M.roughness = q[9];
        M.refraction = q[10];
        M.type = (int)q[11];
    }
    s->nnodes = nnodes;
    s->nodes.assign(bvh_nodes, bvh_nodes + 12 * (size_t)nnodes);
    s->leaf_order.resize(nt);
    orc_bvh_leaf_order(bvh_nodes, nnodes, s->leaf_order.data());
    // Emissive triangles become triangle lights in triangle order
    // (source/amadeus/armada.cu:117-140; emissive test armada.cu:240).
    for (int t = 0; t < nt; t++) {
        const Material &M = s->mats[s->tri_mat[t]];
        if (length(M.emission) > 0.0f) {
            f3 a = s->pos[s->idx[3 * t]], b = s->pos[s->idx[3 * t + 1]], c = s->pos[s->idx[3 * t + 2]];
            s->lights.push_back(TriLight{a, b - a, c - a, M.emission});
        }
    }
    return s;
}

// Material textures of a flat scene; the triangle lights are rebuilt with the
// emissive test of armada.cu:264-265 (|emission| > 0 or an emission texture),
// each with its material's constant emission (armada.cu:143).
void orc_scene_set_textures(orc_scene *s, const float *uvs, const float *tangents, const uint32_t *texels,
                            const uint32_t *dims, int ntex, const int32_t *mat_tex) {
    const size_t nv = s->pos.size(), nm = s->mats.size();
    s->textured = true;
    s->uv.assign(uvs, uvs + 2 * nv);
    if (tangents) s->tan.assign(tangents, tangents + 6 * nv);
    else s->tan.assign(6 * nv, 0.0f);
    s->mat_tex.assign(mat_tex, mat_tex + 5 * nm);
    s->tex_desc.clear();
    size_t total = 0;
    for (int t = 0; t < ntex; t++) {
        s->tex_desc.push_back((uint32_t)total);
        s->tex_desc.push_back(dims[2 * t]);
        s->tex_desc.push_back(dims[2 * t + 1]);
        total += (size_t)dims[2 * t] * dims[2 * t + 1];
    }
    s->texels.assign(texels, texels + total);
    s->lights.clear();
    const int nt = (int)s->tri_mat.size();
    for (int t = 0; t < nt; t++) {
        const uint32_t m = s->tri_mat[t];
        const Material &M = s->mats[m];
        if (length(M.emission) > 0.0f || s->mat_tex[5 * m + 3] >= 0) {
            f3 a = s->pos[s->idx[3 * t]], b = s->pos[s->idx[3 * t + 1]], c = s->pos[s->idx[3 * t + 2]];
            s->lights.push_back(TriLight{a, b - a, c - a, M.emission});
        }
    }
}

// Inverse of a row-major 3x4 affine transform: cofactors of the 3x3 part in
// FP64, translation -R^-1 t, rounded to float once.
void orc_affine_inverse(const float *M, float *Minv) {
    const double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], g = M[8], h = M[9], k = M[10];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double r[9] = {A / det, -(b * k - c * h) / det, (b * f - c * e) / det,
                         B / det, (a * k - c * g) / det, -(a * f - c * d) / det,
                         C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
    const double t[3] = {M[3], M[7], M[11]};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) Minv[4 * i + j] = (float)r[3 * i + j];
        Minv[4 * i + 3] = (float)(-(r[3 * i] * t[0] + r[3 * i + 1] * t[1] + r[3 * i + 2] * t[2]));
    }
}

// World box of an instance: the 8 corners of its mesh's root box through M,
// widened by 1e-4 of the largest extent + 1e-5 of the largest coordinate so
// that rounding of the ray transforms can never cull a hit.
void orc_instance_box(const float *root_min, const float *root_max, const float *M, float *box6) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < 8; k++) {
        f3 p = mk((k & 1) ? root_max[0] : root_min[0], (k & 2) ? root_max[1] : root_min[1],
                  (k & 4) ? root_max[2] : root_min[2]);
        f3 w = xform_point(M, p);
        const float v[3] = {w.x, w.y, w.z};
        for (int a = 0; a < 3; a++) {
            lo[a] = gmin(lo[a], v[a]);
            hi[a] = gmax(hi[a], v[a]);
        }
    }
    float ext = 0.0f, mag = 0.0f;
    for (int a = 0; a < 3; a++) {
        ext = gmax(ext, hi[a] - lo[a]);
        mag = gmax(mag, gmax(fabsf(lo[a]), fabsf(hi[a])));
    }
    const float pad = 1e-4f * ext + 1e-5f * mag;
    for (int a = 0; a < 3; a++) {
        box6[a] = lo[a] - pad;
        box6[3 + a] = hi[a] + pad;
    }
}

orc_scene *orc_scene_create_instanced(int nmesh, const int *mesh_nv, const int *mesh_nt, const float *pos,
                                      const float *nrm, const uint32_t *idx, int ninst, const int *inst_mesh,
                                      const float *inst_xform, const uint32_t *tri_mat, const float *mats, int nm) {
    orc_scene *s = new orc_scene();
    s->instanced = true;
    s->meshes.resize(nmesh);
    size_t vo = 0, to = 0;
    for (int m = 0; m < nmesh; m++) {
        OrcMesh &M = s->meshes[m];
        M.pos.resize(mesh_nv[m]);
        M.nrm.resize(mesh_nv[m]);
        for (int i = 0; i < mesh_nv[m]; i++) {
            M.pos[i] = load3(pos + 3 * (vo + i));
            M.nrm[i] = load3(nrm + 3 * (vo + i));
        }
        M.idx.assign(idx + 3 * to, idx + 3 * (to + mesh_nt[m]));
        std::vector<float> boxes(6 * (size_t)mesh_nt[m]);
        orc_tri_boxes(pos + 3 * vo, M.idx.data(), mesh_nt[m], boxes.data());
        M.nodes.resize(12 * (size_t)(2 * mesh_nt[m] - 1));
        M.nnodes = orc_bvh_build(boxes.data(), mesh_nt[m], M.nodes.data());
        M.leaf_order.resize(mesh_nt[m]);
        orc_bvh_leaf_order(M.nodes.data(), M.nnodes, M.leaf_order.data());
        vo += mesh_nv[m];
        to += mesh_nt[m];
    }
    s->mats.resize(nm);
    for (int m = 0; m < nm; m++) {
        const float *q = mats + 12 * m;
        Material &M = s->mats[m];
        M.diffuse = load3(q);
        M.specular = load3(q + 3);
        M.emission = load3(q + 6);
        M.roughness = q[9];
        M.refraction = q[10];
        M.type = (int)q[11];
    }
    s->inst.resize(ninst);
    std::vector<float> ibox(6 * (size_t)ninst);
    uint32_t base = 0;
    for (int i = 0; i < ninst; i++) {
        OrcInst &I = s->inst[i];
        I.mesh = inst_mesh[i];
        std::memcpy(I.M, inst_xform + 12 * (size_t)i, sizeof(I.M));
        orc_affine_inverse(I.M, I.Minv);
        I.prim_base = base;
        base += (uint32_t)s->meshes[I.mesh].idx.size() / 3;
        const float *root = s->meshes[I.mesh].nodes.data();
        orc_instance_box(root + 4, root + 8, I.M, ibox.data() + 6 * (size_t)i);
    }
    s->tri_mat.assign(tri_mat, tri_mat + base);
    s->tlas.resize(12 * (size_t)(2 * ninst - 1));
    s->ntlas = orc_bvh_build(ibox.data(), ninst, s->tlas.data());
    s->inst_order.resize(ninst);
    orc_bvh_leaf_order(s->tlas.data(), s->ntlas, s->inst_order.data());
    for (int r = 0; r < ninst; r++) s->inst[s->inst_order[r]].rank = r;
    // triangle lights in instance order, then triangle order; vertices through
    // the instance transform (Transform::apply, armada.cu:131-133)
    for (int i = 0; i < ninst; i++) {
        const OrcInst &I = s->inst[i];
        const OrcMesh &m = s->meshes[I.mesh];
        for (size_t t = 0; t < m.idx.size() / 3; t++) {
            const Material &M = s->mats[s->tri_mat[I.prim_base + t]];
            if (length(M.emission) > 0.0f) {
                f3 a = xform_point(I.M, m.pos[m.idx[3 * t]]), b = xform_point(I.M, m.pos[m.idx[3 * t + 1]]),
                   c = xform_point(I.M, m.pos[m.idx[3 * t + 2]]);
                s->lights.push_back(TriLight{a, b - a, c - a, M.emission});
            }
        }
    }
    return s;
}

int orc_scene_tlas(const orc_scene *s, float *nodes) {
    if (nodes) std::memcpy(nodes, s->tlas.data(), sizeof(float) * s->tlas.size());
    return s->ntlas;
}

void orc_scene_destroy(orc_scene *s) { delete s; }

void orc_env_math(const float *in, int n, int which, float *out) {
    for (int i = 0; i < n; i++)
        out[i] = which == 0 ? orc_atan2(in[2 * i], in[2 * i + 1]) : orc_asin(in[2 * i]);
}

void orc_env_lookup(const orc_scene *s, const float *dirs, int n, int flip_v, float *out) {
    const EnvMap env = s->envmap();
    for (int i = 0; i < n; i++) {
        const f3 L = env.tex ? env_radiance(env, load3(dirs + 3 * (size_t)i), flip_v != 0) : mk(0, 0, 0);
        out[3 * (size_t)i] = L.x;
        out[3 * (size_t)i + 1] = L.y;
        out[3 * (size_t)i + 2] = L.z;
    }
}

void orc_tex_lookup(const uint32_t *texels, uint32_t w, uint32_t h, const float *uv, int n, float *out) {
    Texture t;
    t.texels = texels;
    t.w = w;
    t.h = h;
    for (int i = 0; i < n; i++) tex_fetch(t, uv[2 * (size_t)i], uv[2 * (size_t)i + 1], out + 4 * (size_t)i);
}

void orc_env_sample_dirs(uint32_t seed_word, int n, float *out) {
    f3 seed = seed_from_words(seed_word, 0x51u, 0x7u);
    for (int i = 0; i < n; i++) {
        const f3 d = env_sample_dir(seed);
        out[3 * (size_t)i] = d.x;
        out[3 * (size_t)i + 1] = d.y;
        out[3 * (size_t)i + 2] = d.z;
    }
}

void orc_scene_set_envmap(orc_scene *s, const float *rgba, uint32_t w, uint32_t h) {
    if (!rgba || w == 0 || h == 0) {
        s->env.clear();
        s->env_w = s->env_h = 0;
        return;
    }
    s->env.assign(rgba, rgba + 4 * (size_t)w * h);
    s->env_w = w;
    s->env_h = h;
}
int orc_scene_light_count(const orc_scene *s) { return (int)s->lights.size(); }
void orc_scene_lights(const orc_scene *s, float *out) {
    for (size_t i = 0; i < s->lights.size(); i++) {
        const TriLight &l = s->lights[i];
        f3 v[4] = {l.a, l.ab, l.ac, l.intensity};
        for (int k = 0; k < 4; k++) {
            out[12 * i + 3 * k] = v[k].x;
            out[12 * i + 3 * k + 1] = v[k].y;
            out[12 * i + 3 * k + 2] = v[k].z;
        }
    }
}

// ---------------------------------------------------------------------------
// RNG probes
// ---------------------------------------------------------------------------
void orc_pcg3d(const uint32_t *in, uint32_t *out, int n) {
    for (int i = 0; i < n; i++) {
        uint32_t v[3] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        pcg3d(v);
        out[3 * i] = v[0];
        out[3 * i + 1] = v[1];
        out[3 * i + 2] = v[2];
    }
}

void orc_pcg3f(float *seeds, int n) {
    for (int i = 0; i < n; i++) {
        f3 s = load3(seeds + 3 * i);
        pcg3f(s);
        seeds[3 * i] = s.x;
        seeds[3 * i + 1] = s.y;
        seeds[3 * i + 2] = s.z;
    }
}

void orc_sincos_turn(const float *u, float *c, float *s, int n) {
    for (int i = 0; i < n; i++) sincos_turn(u[i], c[i], s[i]);
}

// ---------------------------------------------------------------------------
// BVH build — source/bvh.cpp restated iteratively
// ---------------------------------------------------------------------------
struct Box {
    float mn[3], mx[3];
};

// include/mesh.hpp:66-71: min(p1, min(p2, p3)), max likewise (glm semantics)
void orc_tri_boxes(const float *pos, const uint32_t *idx, int nt, float *boxes) {
    for (int t = 0; t < nt; t++) {
        const float *p1 = pos + 3 * idx[3 * t], *p2 = pos + 3 * idx[3 * t + 1], *p3 = pos + 3 * idx[3 * t + 2];
        for (int k = 0; k < 3; k++) {
            boxes[6 * t + k] = gmin(p1[k], gmin(p2[k], p3[k]));
            boxes[6 * t + 3 + k] = gmax(p1[k], gmax(p2[k], p3[k]));
        }
    }
}

namespace {

struct BNode {
    Box box;
    int object = -1;
    int left = -1, right = -1;
};

// include/bbox.hpp:16-26
inline float surface_area(const float mn[3], const float mx[3]) {
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    float xy = dx * dy, yz = dy * dz, xz = dx * dz;
    return 2.0f * (xy + yz + xz);
}

// bvh.cpp:32-43 (fold from the first element)
Box union_of(const std::vector<BNode> &N, const std::vector<int> &items) {
    Box b = N[items[0]].box;
    for (size_t i = 1; i < items.size(); i++)
        for (int k = 0; k < 3; k++) {
            b.mn[k] = gmin(b.mn[k], N[items[i]].box.mn[k]);
            b.mx[k] = gmax(b.mx[k], N[items[i]].box.mx[k]);
        }
    return b;
}

}  // namespace

int orc_bvh_build(const float *boxes, int n, float *out) {
    if (n <= 0) return 0;
    std::vector<BNode> N(n);
    N.reserve(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            N[i].box.mn[k] = boxes[6 * i + k];
            N[i].box.mx[k] = boxes[6 * i + 3 + k];
        }
        N[i].object = i;  // bvh.cpp:198-207
    }
    struct Task {
        std::vector<int> items;
        int parent, side;
    };
    int root = -1;
    std::vector<Task> stack;
    {
        Task t;
        t.items.resize(n);
        for (int i = 0; i < n; i++) t.items[i] = i;
        t.parent = -1;
        t.side = 0;
        stack.push_back(std::move(t));
    }
    auto link = [&](int parent, int side, int node) {
        if (parent < 0) root = node;
        else if (side == 0) N[parent].left = node;
        else N[parent].right = node;
    };
    while (!stack.empty()) {
        Task task = std::move(stack.back());
        stack.pop_back();
        const std::vector<int> &items = task.items;
        size_t cnt = items.size();
        if (cnt == 1) {  // bvh.cpp:108-109
            link(task.parent, task.side, items[0]);
            continue;
        }
        if (cnt == 2) {  // bvh.cpp:111-117
            BNode node;
            node.box = union_of(N, items);
            node.left = items[0];
            node.right = items[1];
            N.push_back(node);
            link(task.parent, task.side, (int)N.size() - 1);
            continue;
        }
        // Axis and interval of the single box with the largest extent (bvh.cpp:119-139)
        int axis = 0;
        float max_extent = 0.0f, min_value = FLT_MAX, max_value = -FLT_MAX;
        for (size_t j = 0; j < cnt; j++) {
            const Box &b = N[items[j]].box;
            for (int i = 0; i < 3; i++) {
                float extent = fabsf(b.mx[i] - b.mn[i]);
                if (extent > max_extent) {
                    max_extent = extent;
                    min_value = b.mn[i];
                    max_value = b.mx[i];
                    axis = i;
                }
            }
        }
        // 10 candidate planes; sah_cost (bvh.cpp:46-98) evaluated for all ten in
        // one pass — each candidate's fold visits the items in the same order.
        const int bins = 10;
        float split[bins];
        for (int i = 0; i < bins; i++) split[i] = (max_value - min_value) / (float)bins * (float)i + min_value;
        float lmn[bins][3], lmx[bins][3], rmn[bins][3], rmx[bins][3];
        int pl[bins], pr[bins];
        for (int i = 0; i < bins; i++) {
            pl[i] = pr[i] = 0;
            for (int k = 0; k < 3; k++) {
                lmn[i][k] = rmn[i][k] = FLT_MAX;
                lmx[i][k] = rmx[i][k] = -FLT_MAX;
            }
        }
        float tmn[3], tmx[3];
        for (int k = 0; k < 3; k++) {
            tmn[k] = N[items[0]].box.mn[k];
            tmx[k] = N[items[0]].box.mx[k];
        }
        for (size_t j = 0; j < cnt; j++) {
            const Box &b = N[items[j]].box;
            for (int k = 0; k < 3; k++) {
                tmn[k] = gmin(tmn[k], b.mn[k]);
                tmx[k] = gmax(tmx[k], b.mx[k]);
            }
            float value = (b.mn[axis] + b.mx[axis]) / 2.0f;
            for (int i = 0; i < bins; i++) {
                if (value < split[i]) {
                    pl[i]++;
                    for (int k = 0; k < 3; k++) {
                        lmn[i][k] = gmin(lmn[i][k], b.mn[k]);
                        lmx[i][k] = gmax(lmx[i][k], b.mx[k]);
                    }
                } else {
                    pr[i]++;
                    for (int k = 0; k < 3; k++) {
                        rmn[i][k] = gmin(rmn[i][k], b.mn[k]);
                        rmx[i][k] = gmax(rmx[i][k], b.mx[k]);
                    }
                }
            }
        }
        float sa_total = surface_area(tmn, tmx);
        float min_cost = FLT_MAX, min_split = 0.0f;
        for (int i = 0; i < bins; i++) {
            float cost;
            if (pl[i] == 0 || pr[i] == 0) cost = FLT_MAX;
            else
                cost = 1.0f + ((float)pl[i] * surface_area(lmn[i], lmx[i]) +
                               (float)pr[i] * surface_area(rmn[i], rmx[i])) / sa_total;
            if (cost < min_cost) {  // bvh.cpp:150-153 (strict: first minimum)
                min_cost = cost;
                min_split = split[i];
            }
        }
        Task L, R;
        if (min_cost == FLT_MAX) {  // bvh.cpp:159-166 even/odd fallback
            for (size_t j = 0; j < cnt; j++) (j % 2 == 0 ? L : R).items.push_back(items[j]);
        } else {  // bvh.cpp:167-180
            for (size_t j = 0; j < cnt; j++) {
                const Box &b = N[items[j]].box;
                float value = (b.mn[axis] + b.mx[axis]) / 2.0f;
                (value < min_split ? L : R).items.push_back(items[j]);
            }
        }
        BNode node;
        node.box = union_of(N, items);
        N.push_back(node);
        int id = (int)N.size() - 1;
        link(task.parent, task.side, id);
        L.parent = R.parent = id;
        L.side = 0;
        R.side = 1;
        stack.push_back(std::move(R));
        stack.push_back(std::move(L));
    }
    // subtree node counts (bvh.cpp:18-22), post-order without recursion
    std::vector<int> count(N.size(), 1);
    {
        std::vector<std::pair<int, int>> st;
        st.push_back({root, 0});
        while (!st.empty()) {
            auto &top = st.back();
            int v = top.first;
            if (N[v].left < 0) {
                st.pop_back();
                continue;
            }
            if (top.second == 0) {
                top.second = 1;
                st.push_back({N[v].left, 0});
            } else if (top.second == 1) {
                top.second = 2;
                st.push_back({N[v].right, 0});
            } else {
                count[v] = 1 + count[N[v].left] + count[N[v].right];
                st.pop_back();
            }
        }
    }
    // serialize (bvh.cpp:213-260): preorder, 3 vec4 per node, links in vec4 units
    std::vector<std::pair<int, int>> st;  // (node, miss)
    st.push_back({root, -1});
    int written = 0;
    while (!st.empty()) {
        auto [v, miss] = st.back();
        st.pop_back();
        int size = 3 * written;
        int after = size + 3;
        int left = N[v].left, right = N[v].right;
        int left_index = left >= 0 ? after : -1;
        int left_size = left >= 0 ? 3 * count[left] : 0;
        int right_index = right >= 0 ? after + left_size : -1;
        int hit = left_index;
        if (left < 0 && right < 0) hit = miss;
        int miss_left = right >= 0 ? right_index : miss;
        float *o = out + 12 * (size_t)written;
        o[0] = N[v].object != -1 ? 1.0f : 0.0f;
        o[1] = bitsf((uint32_t)N[v].object);
        o[2] = bitsf((uint32_t)hit);
        o[3] = bitsf((uint32_t)miss);
        for (int k = 0; k < 3; k++) {
            o[4 + k] = N[v].box.mn[k];
            o[8 + k] = N[v].box.mx[k];
        }
        o[7] = 0.0f;
        o[11] = 0.0f;
        written++;
        if (right >= 0) st.push_back({right, miss});
        if (left >= 0) st.push_back({left, miss_left});
    }
    return written;
}

int orc_bvh_leaf_order(const float *nodes, int nnodes, int32_t *order) {
    int k = 0;
    for (int i = 0; i < nnodes; i++) {
        int32_t obj = (int32_t)fbits(nodes[12 * i + 1]);
        if (obj != -1) order[k++] = obj;
    }
    return k;
}

int orc_bvh_depth(const float *nodes, int nnodes) {
    // depth via preorder walk: a node's children follow it; track with a stack of
    // remaining right-link targets.
    std::vector<std::pair<int, int>> st;  // (vec4 index, depth)
    if (nnodes == 0) return 0;
    st.push_back({0, 1});
    int best = 0;
    while (!st.empty()) {
        auto [v, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        const float *h = nodes + 4 * (size_t)v;
        int32_t obj = (int32_t)fbits(h[1]);
        if (obj != -1) continue;
        int left = (int32_t)fbits(h[2]);
        int right = (int32_t)fbits(nodes[4 * (size_t)left + 3]);
        st.push_back({left, d + 1});
        st.push_back({right, d + 1});
    }
    return best;
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
namespace {

// Triangles + bvh.cpp tree of one mesh (or of a flat scene).
struct Geo {
    const f3 *pos;
    const uint32_t *idx;
    const float *nodes;
    int nnodes;
    const int32_t *leaf_order;
    int nt;
};

inline Geo geo_of(const orc_scene *s) {
    return Geo{s->pos.data(), s->idx.data(), s->nodes.data(), s->nnodes, s->leaf_order.data(),
               (int)s->leaf_order.size()};
}
inline Geo geo_of(const OrcMesh &m) {
    return Geo{m.pos.data(), m.idx.data(), m.nodes.data(), m.nnodes, m.leaf_order.data(), (int)m.leaf_order.size()};
}

inline void tri_verts(const Geo &g, int prim, f3 &a, f3 &b, f3 &c) {
    a = g.pos[g.idx[3 * prim]];
    b = g.pos[g.idx[3 * prim + 1]];
    c = g.pos[g.idx[3 * prim + 2]];
}
inline void tri_verts(const orc_scene *s, int prim, f3 &a, f3 &b, f3 &c) { tri_verts(geo_of(s), prim, a, b, c); }

// shaders/rt/modules/bbox.glsl:8-57 (returns t entry or -1)
inline float glsl_box(f3 o, f3 d, const float *mn, const float *mx) {
    float tmin = (mn[0] - o.x) / d.x, tmax = (mx[0] - o.x) / d.x;
    if (tmin > tmax) std::swap(tmin, tmax);
    float tymin = (mn[1] - o.y) / d.y, tymax = (mx[1] - o.y) / d.y;
    if (tymin > tymax) std::swap(tymin, tymax);
    if ((tmin > tymax) || (tymin > tmax)) return -1.0f;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (mn[2] - o.z) / d.z, tzmax = (mx[2] - o.z) / d.z;
    if (tzmin > tzmax) std::swap(tzmin, tzmax);
    if ((tmin > tzmax) || (tzmin > tmax)) return -1.0f;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return tmin;
}

// bbox.glsl:60-66
inline bool in_box(f3 p, const float *mn, const float *mx) {
    return !(p.x < mn[0] || p.y < mn[1] || p.z < mn[2] || p.x > mx[0] || p.y > mx[1] || p.z > mx[2]);
}

// Conservative slab test: never rejects a box that the ray's segment [0, tlim]
// meets (a zero direction component constrains only by the origin, a relative
// slack absorbs rounding).  Over-acceptance only costs visits.
inline bool robust_box(f3 o, f3 d, const float *mn, const float *mx, float tlim) {
    float tn = -FLT_MAX, tf = FLT_MAX;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; a++) {
        if (dd[a] == 0.0f) {
            if (oo[a] < mn[a] || oo[a] > mx[a]) return false;
            continue;
        }
        float t0 = (mn[a] - oo[a]) / dd[a], t1 = (mx[a] - oo[a]) / dd[a];
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0);
        tf = std::min(tf, t1);
    }
    float eps = 1e-4f * std::max(fabsf(tn), fabsf(tf)) + 1e-6f;
    return tn - eps <= tf + eps && tf + eps >= 0.0f && tn - eps <= tlim;
}

// Closest hit: lexicographic minimum of (t, preorder leaf rank) over triangles
// with 0 < t < tmax.  mode 0 brute force in rank order, 1 GLSL walk, 2 robust walk.
Hit closest_geo(const Geo &s, f3 o, f3 d, float tmax, int mode, Counters &c) {
    Hit h{0.0f, 0.0f, 0.0f, -1};
    float best = tmax;
    if (mode == 0) {
        for (int k = 0; k < s.nt; k++) {
            const int32_t prim = s.leaf_order[k];
            f3 a, b, cc;
            tri_verts(s, prim, a, b, cc);
            float b1 = 0.0f, b2 = 0.0f;
            float t = intersect_tri(a, b, cc, o, d, b1, b2);
            c.tris++;
            if (t > 0.0f && t < best) {
                best = t;
                h = Hit{t, b1, b2, prim};
            }
        }
        return h;
    }
    const float *nodes = s.nodes;
    int node = s.nnodes > 0 ? 0 : -1;
    while (node != -1) {
        const float *hd = nodes + 4 * (size_t)node;
        int32_t obj = (int32_t)fbits(hd[1]);
        c.nodes++;
        if (obj != -1) {
            f3 a, b, cc;
            tri_verts(s, obj, a, b, cc);
            float b1 = 0.0f, b2 = 0.0f;
            float t = intersect_tri(a, b, cc, o, d, b1, b2);
            c.tris++;
            if (t > 0.0f && t < best) {
                best = t;
                h = Hit{t, b1, b2, obj};
            }
            node = (int32_t)fbits(hd[3]);
        } else {
            const float *mn = hd + 4, *mx = hd + 8;
            bool go;
            if (mode == 1) {
                float t = glsl_box(o, d, mn, mx);
                go = (t > 0.0f && t < best) || in_box(o, mn, mx);
            } else {
                go = robust_box(o, d, mn, mx, best);
            }
            node = go ? (int32_t)fbits(hd[2]) : (int32_t)fbits(hd[3]);
        }
    }
    return h;
}

// Any hit with 0 < t < tmax (OPTIX_RAY_FLAG_TERMINATE_ON_FIRST_HIT,
// amadeus_common.cuh:190-212).
bool occluded_geo(const Geo &s, f3 o, f3 d, float tmax, Counters &c) {
    const float *nodes = s.nodes;
    int node = s.nnodes > 0 ? 0 : -1;
    while (node != -1) {
        const float *hd = nodes + 4 * (size_t)node;
        int32_t obj = (int32_t)fbits(hd[1]);
        c.nodes++;
        if (obj != -1) {
            f3 a, b, cc;
            tri_verts(s, obj, a, b, cc);
            float b1, b2;
            float t = intersect_tri(a, b, cc, o, d, b1, b2);
            c.tris++;
            if (t > 0.0f && t < tmax) return true;
            node = (int32_t)fbits(hd[3]);
        } else {
            bool go = robust_box(o, d, hd + 4, hd + 8, tmax);
            node = go ? (int32_t)fbits(hd[2]) : (int32_t)fbits(hd[3]);
        }
    }
    return false;
}

// Two-level closest hit: the world ray walks the instance tree; at an
// instance it is taken into the mesh's space by the inverse transform (t is
// unchanged: the direction is not renormalised) and walks the mesh tree with
// the best t so far.  Ties keep the first instance in instance-tree preorder,
// then the first triangle in mesh preorder: the lexicographic minimum of
// (t, instance rank, triangle rank).
Hit closest_inst(const orc_scene *s, f3 o, f3 d, float tmax, int mode, Counters &c) {
    Hit h{0.0f, 0.0f, 0.0f, -1};
    float best = tmax;
    auto visit = [&](int i) {
        const OrcInst &I = s->inst[i];
        Hit hm = closest_geo(geo_of(s->meshes[I.mesh]), xform_point(I.Minv, o), xform_dir(I.Minv, d), best,
                             mode == 0 ? 0 : 2, c);
        if (hm.prim >= 0) {
            best = hm.t;
            h = hm;
            h.prim = (int)(I.prim_base + (uint32_t)hm.prim);
        }
    };
    if (mode == 0) {
        for (int32_t i : s->inst_order) visit(i);
        return h;
    }
    const float *nodes = s->tlas.data();
    int node = s->ntlas > 0 ? 0 : -1;
    while (node != -1) {
        const float *hd = nodes + 4 * (size_t)node;
        int32_t obj = (int32_t)fbits(hd[1]);
        c.nodes++;
        if (obj != -1) {
            visit(obj);
            node = (int32_t)fbits(hd[3]);
        } else {
            bool go = robust_box(o, d, hd + 4, hd + 8, best);
            node = go ? (int32_t)fbits(hd[2]) : (int32_t)fbits(hd[3]);
        }
    }
    return h;
}

bool occluded_inst(const orc_scene *s, f3 o, f3 d, float tmax, Counters &c) {
    const float *nodes = s->tlas.data();
    int node = s->ntlas > 0 ? 0 : -1;
    while (node != -1) {
        const float *hd = nodes + 4 * (size_t)node;
        int32_t obj = (int32_t)fbits(hd[1]);
        c.nodes++;
        if (obj != -1) {
            const OrcInst &I = s->inst[obj];
            if (occluded_geo(geo_of(s->meshes[I.mesh]), xform_point(I.Minv, o), xform_dir(I.Minv, d), tmax, c))
                return true;
            node = (int32_t)fbits(hd[3]);
        } else {
            bool go = robust_box(o, d, hd + 4, hd + 8, tmax);
            node = go ? (int32_t)fbits(hd[2]) : (int32_t)fbits(hd[3]);
        }
    }
    return false;
}

Hit closest(const orc_scene *s, f3 o, f3 d, float tmax, int mode, Counters &c) {
    return s->instanced ? closest_inst(s, o, d, tmax, mode, c) : closest_geo(geo_of(s), o, d, tmax, mode, c);
}
bool occluded(const orc_scene *s, f3 o, f3 d, float tmax, Counters &c) {
    return s->instanced ? occluded_inst(s, o, d, tmax, c) : occluded_geo(geo_of(s), o, d, tmax, c);
}

// Instance of a global triangle id (instances hold consecutive id ranges).
inline int instance_of(const orc_scene *s, int prim) {
    int lo = 0, hi = (int)s->inst.size() - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (s->inst[mid].prim_base <= (uint32_t)prim) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

void orc_trace(const orc_scene *s, const float *rays, int n, int mode, float *hits, uint64_t *counters) {
    Counters c;
    for (int i = 0; i < n; i++) {
        const float *r = rays + 8 * (size_t)i;
        Hit h = closest(s, load3(r), load3(r + 4), r[3], mode, c);
        hits[4 * i] = h.t;
        hits[4 * i + 1] = h.b1;
        hits[4 * i + 2] = h.b2;
        hits[4 * i + 3] = bitsf((uint32_t)h.prim);
    }
    if (counters) {
        counters[0] += c.nodes;
        counters[1] += c.tris;
    }
}

void orc_occluded(const orc_scene *s, const float *rays, int n, int32_t *out) {
    Counters c;
    for (int i = 0; i < n; i++) {
        const float *r = rays + 8 * (size_t)i;
        out[i] = occluded(s, load3(r), load3(r + 4), r[3], c) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// BSDF probes
// ---------------------------------------------------------------------------
static Material probe_mat(const float *kd, const float *ks, float rough) {
    Material m;
    m.diffuse = load3(kd);
    m.specular = load3(ks);
    m.emission = mk(0, 0, 0);
    m.roughness = rough;
    m.refraction = 1.0f;
    m.type = eDiffuse;
    return m;
}

float orc_ggx_pdf(const float *n, const float *wi, const float *wo, const float *kd, const float *ks,
                  float roughness) {
    return ggx_pdf(probe_mat(kd, ks, roughness), load3(n), load3(wi), load3(wo));
}

void orc_ggx_brdf(const float *n, const float *wi, const float *wo, const float *kd, const float *ks,
                  float roughness, float *out) {
    SurfaceHit sh;
    sh.mat = probe_mat(kd, ks, roughness);
    sh.n = load3(n);
    sh.wo = load3(wo);
    sh.entering = true;
    f3 f = brdf(sh, load3(wi), eDiffuse);
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
}

void orc_ggx_sample(const float *n, const float *wo, const float *kd, const float *ks, float roughness,
                    float *seed, float *wi_out) {
    f3 sd = load3(seed);
    f3 wi = ggx_sample(probe_mat(kd, ks, roughness), load3(n), load3(wo), sd);
    seed[0] = sd.x;
    seed[1] = sd.y;
    seed[2] = sd.z;
    wi_out[0] = wi.x;
    wi_out[1] = wi.y;
    wi_out[2] = wi.z;
}

void orc_ggx_sample_batch(const float *n, const float *wo, const float *kd, const float *ks, float roughness,
                          int count, uint32_t word, float *wi_out) {
    const Material m = probe_mat(kd, ks, roughness);
    for (int i = 0; i < count; i++) {
        f3 sd = seed_from_words((uint32_t)i, word, 0x51u);
        f3 wi = ggx_sample(m, load3(n), load3(wo), sd);
        wi_out[3 * i] = wi.x;
        wi_out[3 * i + 1] = wi.y;
        wi_out[3 * i + 2] = wi.z;
    }
}

void orc_glass_sample(const float *n, const float *wi, float eta, int entering, float *out) {
    SurfaceHit sh;
    sh.mat = probe_mat(n, n, 1.0f);
    sh.mat.refraction = eta;
    sh.mat.type = eTransmission;
    sh.n = load3(n);
    sh.entering = entering != 0;
    f3 d = glass_sample(sh, load3(wi));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

void orc_light_sample(const float *l, float *seed, float *point_out) {
    TriLight L{load3(l), load3(l + 3), load3(l + 6), load3(l + 9)};
    f3 sd = load3(seed);
    f3 p = sample_area_light(L, sd);
    seed[0] = sd.x;
    seed[1] = sd.y;
    seed[2] = sd.z;
    point_out[0] = p.x;
    point_out[1] = p.y;
    point_out[2] = p.z;
}

// ---------------------------------------------------------------------------
// Shared per-hit logic (amadeus_common.cuh)
// ---------------------------------------------------------------------------
namespace {

const float kEps = 1e-3f;          // amadeus_common.cuh:22
const float kShadowEps = 0.05f;    // amadeus_common.cuh:193
const float kTraceTmax = 1e16f;    // amadeus_common.cuh:472

struct Ctx {
    const orc_scene *s;
    Counters c;
};

// LOAD_INTERSECTION_DATA (amadeus_common.cuh:490-515), calculate_normal
// (:74-112), calculate_intersection (:62-71) for world-space triangles.
struct Surface {
    f3 x, n;
    bool entering;
    Material mat;
    float u = 0.0f, v = 0.0f;  // texture coordinates (textured scenes)
};

inline f3 interp(f3 a, f3 b, f3 c, float b1, float b2) {
    float w = 1.0f - b1 - b2;
    return (w * a + b1 * b) + b2 * c;
}

// Instanced hit: calculate_intersection / calculate_normal with the
// instance's model matrix (amadeus_common.cuh:62-112): x = M (interp, 1),
// edges and the interpolated normal by M (w = 0).
bool load_surface_inst(const orc_scene *s, const Hit &h, f3 dir, Surface &sf) {
    const OrcInst &I = s->inst[instance_of(s, h.prim)];
    const OrcMesh &m = s->meshes[I.mesh];
    const int prim = h.prim - (int)I.prim_base;
    const uint32_t *ix = &m.idx[3 * prim];
    f3 a = m.pos[ix[0]], b = m.pos[ix[1]], c = m.pos[ix[2]];
    f3 e1 = xform_dir(I.M, b - a), e2 = xform_dir(I.M, c - a);
    f3 ng = normalize(cross(e1, e2));
    if (dot(ng, dir) > 0.0f) {
        ng = -ng;
        sf.entering = false;
    } else {
        sf.entering = true;
    }
    ng = normalize(ng);
    f3 nn = xform_dir(I.M, interp(m.nrm[ix[0]], m.nrm[ix[1]], m.nrm[ix[2]], h.b1, h.b2));
    if (dot(nn, ng) < 0.0f) nn = -nn;
    sf.n = normalize(nn);
    sf.x = xform_point(I.M, interp(a, b, c, h.b1, h.b2));
    sf.mat = s->mats[s->tri_mat[h.prim]];
    return !isnan3(sf.n);
}

// normal_map = false: the editor's ray-cast G-buffer (gbuffer_rtx_shader.cu:
// 95-160), which records (u, 1 - v) but neither textures nor normal maps.
bool load_surface(const orc_scene *s, const Hit &h, f3 dir, Surface &sf, bool normal_map = true) {
    if (s->instanced) return load_surface_inst(s, h, dir, sf);
    f3 a, b, c;
    tri_verts(s, h.prim, a, b, c);
    f3 e1 = b - a, e2 = c - a;
    f3 ng = normalize(cross(e1, e2));
    if (dot(ng, dir) > 0.0f) {
        ng = -ng;
        sf.entering = false;
    } else {
        sf.entering = true;
    }
    ng = normalize(ng);
    const uint32_t *ix = &s->idx[3 * h.prim];
    f3 nn = interp(s->nrm[ix[0]], s->nrm[ix[1]], s->nrm[ix[2]], h.b1, h.b2);
    if (dot(nn, ng) < 0.0f) nn = -nn;
    sf.n = normalize(nn);
    sf.x = interp(a, b, c, h.b1, h.b2);
    const uint32_t m = s->tri_mat[h.prim];
    sf.mat = s->mats[m];
    if (s->textured) {
        // uv: interpolated tex_coords, v flipped (amadeus_common.cuh:498-502)
        const float *ta = &s->uv[2 * (size_t)ix[0]], *tb = &s->uv[2 * (size_t)ix[1]], *tc = &s->uv[2 * (size_t)ix[2]];
        const float w = 1.0f - h.b1 - h.b2;
        const float u = (w * ta[0] + h.b1 * tb[0]) + h.b2 * tc[0];
        const float v = 1.0f - ((w * ta[1] + h.b1 * tb[1]) + h.b2 * tc[1]);
        sf.u = u;
        sf.v = v;
        if (!normal_map) return !isnan3(sf.n);
        const int32_t *mt = &s->mat_tex[5 * (size_t)m];
        float t4[4];
        // calculate_material (:154-187)
        if (mt[0] >= 0) {
            tex_fetch(s->texture(mt[0]), u, v, t4);
            sf.mat.diffuse = mk(t4[0], t4[1], t4[2]);
        }
        if (mt[2] >= 0) {
            tex_fetch(s->texture(mt[2]), u, v, t4);
            sf.mat.specular = mk(t4[0], t4[1], t4[2]);
        }
        if (mt[3] >= 0) {
            tex_fetch(s->texture(mt[3]), u, v, t4);
            sf.mat.emission = mk(t4[0], t4[1], t4[2]);
        }
        if (mt[4] >= 0) {
            tex_fetch(s->texture(mt[4]), u, v, t4);
            sf.mat.roughness = t4[0];
        }
        // calculate_normal's normal map (:114-146): tbn = mat3(normalize(T),
        // normalize(B), normalize(n)) with T and B normalised once before
        // (glm::normalize(model * ...)); n = normalize(tbn * (2 rgb - 1))
        if (mt[1] >= 0) {
            tex_fetch(s->texture(mt[1]), u, v, t4);
            const f3 tn = mk(2.0f * t4[0] - 1.0f, 2.0f * t4[1] - 1.0f, 2.0f * t4[2] - 1.0f);
            const float *A = &s->tan[6 * (size_t)ix[0]], *B = &s->tan[6 * (size_t)ix[1]], *C = &s->tan[6 * (size_t)ix[2]];
            const f3 T = normalize(normalize(interp(load3(A), load3(B), load3(C), h.b1, h.b2)));
            const f3 Bt = normalize(normalize(interp(load3(A + 3), load3(B + 3), load3(C + 3), h.b1, h.b2)));
            const f3 N = normalize(sf.n);
            const f3 col0 = T * tn.x, col1 = Bt * tn.y, col2 = N * tn.z;
            sf.n = normalize((col0 + col1) + col2);
        }
    }
    return !isnan3(sf.n);
}

inline unsigned pick_light(f3 &seed, unsigned total) {
    unsigned i = (unsigned)(rand_uniform(seed) * (float)total);
    return i < total ? i : total - 1;  // rand*total may round up to total
}

// Ld (amadeus_common.cuh:255-286) with Ld_light (lighting.cuh:149-174) and
// Ld_Environment (amadeus_common.cuh:215-252) when the environment map is on
f3 Ld(Ctx &cx, const SurfaceHit &sh, f3 &seed) {
    const orc_scene *s = cx.s;
    const EnvMap env = s->envmap();
    const int nl = (int)s->lights.size();
    int total = nl + (env.tex ? 1 : 0);
    if (total == 0) {
        rand_uniform(seed);
        return mk(0, 0, 0);
    }
    unsigned i = pick_light(seed, (unsigned)total);
    float pdf = 1.0f / (float)total;
    if ((int)i >= nl) {
        // Ld_Environment: a direction about world +z, pdf 1/(4 pi), the
        // shadow ray to WORLD_RADIUS
        f3 wi = env_sample_dir(seed);
        f3 Li = env_radiance(env, wi, false);
        float light_pdf = kEnvPdf;
        cx.c.shadow++;
        f3 contr;
        if (occluded(s, sh.x, wi, kEnvDistance - kShadowEps, cx.c)) {
            contr = mk(0, 0, 0);
        } else {
            f3 f = brdf(sh, wi, eDiffuse) * fabsf(dot(sh.n, wi));
            contr = f * Li;
        }
        pdf *= light_pdf;
        return contr / pdf;
    }
    const TriLight &light = s->lights[i];
    float light_pdf = 1.0f / light_area(light);
    f3 lpos = sample_area_light(light, seed);
    f3 wi = normalize(lpos - sh.x);
    float R = length(lpos - sh.x);
    cx.c.shadow++;
    f3 contr;
    if (occluded(s, sh.x, wi, R - kShadowEps, cx.c)) {
        contr = mk(0, 0, 0);
    } else {
        f3 f = brdf(sh, wi, eDiffuse);
        float ldot = fabsf(dot(light_normal(light), wi));
        float geometric = ldot * fabsf(dot(sh.n, wi)) / (R * R);
        contr = f * light.intensity * geometric;
    }
    pdf *= light_pdf;
    return contr / pdf;
}

struct RayPacket {  // amadeus_path_tracer.cu:10-23
    f3 value{0, 0, 0};
    f3 position{0, 0, 0};
    f3 normal{0, 0, 0};
    f3 albedo{0, 0, 0};
    f3 seed{0, 0, 0};
    int depth = 0;
    float t = 0.0f;  // hit distance of the outermost level (ReSTIR GI sample point)
};

// __closesthit__ / __miss__ of amadeus_path_tracer.cu:133-220, recursive as in
// the reference.  The ray past max_depth is not traced: its program only
// writes value = 0 (LOAD_RAYPACKET, amadeus_common.cuh:480-488).
void pt_trace(Ctx &cx, int max_depth, f3 o, f3 d, RayPacket &rp) {
    if (rp.depth > max_depth) {
        cx.c.discarded++;
        rp.value = mk(0, 0, 0);
        return;
    }
    cx.c.closest++;
    Hit h = closest(cx.s, o, d, kTraceTmax, 2, cx.c);
    if (h.prim < 0) {  // __miss__: the environment map at (u, 1 - v), or 0
        const EnvMap env = cx.s->envmap();
        rp.value = env.tex ? env_radiance(env, d, true) : mk(0, 0, 0);
        return;
    }
    Surface sf;
    if (!load_surface(cx.s, h, d, sf)) {
        rp.value = mk(0, 0, 0);
        return;
    }
    const Material &material = sf.mat;
    f3 n = sf.n;
    f3 x = sf.x + ((material.type == eTransmission ? -1.0f : 1.0f) * n) * kEps;
    SurfaceHit sh{material, sf.entering, n, -d, x};
    f3 direct = material.emission + Ld(cx, sh, rp.seed);
    int out;
    f3 wi;
    float pdf;
    f3 f = eval_bsdf(sh, wi, pdf, out, rp.seed);
    f3 T = f * fabsf(dot(wi, n)) / pdf;
    rp.depth++;
    f3 indirect = mk(0, 0, 0);
    if (pdf > 0.0f) {
        pt_trace(cx, max_depth, x, wi, rp);
        indirect = rp.value;
    }
    rp.value = direct;
    if (pdf > 0.0f) rp.value = rp.value + T * indirect;
    rp.position = x;
    rp.normal = n;
    rp.albedo = material.diffuse;
    rp.t = h.t;
}

// make_ray (amadeus_path_tracer.cu:26-56; ReSTIR variant amadeus_restir.cu:31-65
// uses a fixed -0.5 offset).  The seed is advanced once by pcg3f either way.
void make_ray(const orc_camera *cam, const orc_params *p, uint32_t px, uint32_t py, f3 &seed, bool jitter,
              f3 &o, f3 &d) {
    pcg3f(seed);
    float xoff = jitter ? fract(seed.x) - 0.5f : -0.5f;
    float yoff = jitter ? fract(seed.y) - 0.5f : -0.5f;
    float dx = 2.0f * (((float)px + xoff) / (float)p->width) - 1.0f;
    float dy = 2.0f * (((float)py + yoff) / (float)p->height) - 1.0f;
    o = load3(cam->center);
    d = normalize((dx * load3(cam->U) + dy * load3(cam->V)) + load3(cam->W));
}

// accumulate (amadeus_path_tracer.cu:59-70) on a glm::vec4 (true division)
inline void accumulate4(float *dst, const float *sample, const orc_params *p) {
    if (p->accumulate) {
        float count = (float)p->samples;
        for (int k = 0; k < 4; k++) dst[k] = (dst[k] * count + sample[k]) / (count + 1.0f);
    } else {
        for (int k = 0; k < 4; k++) dst[k] = sample[k];
    }
}

void pt_sample(Ctx &cx, const orc_camera *cam, const orc_params *p, uint32_t px, uint32_t py, RayPacket &rp) {
    rp = RayPacket();
    rp.seed = pixel_seed(px, py, (uint32_t)p->samples, p->user_seed);
    f3 o, d;
    make_ray(cam, p, px, py, rp.seed, p->fixed_offset == 0, o, d);
    pt_trace(cx, p->max_depth, o, d, rp);
}

template <class F>
void parallel_rows(int height, int nthreads, std::vector<Counters> &per, F fn) {
    if (nthreads < 1) nthreads = 1;
    per.assign(nthreads, Counters());
    std::atomic<int> next{0};
    auto work = [&](int tid) {
        for (;;) {
            int y = next.fetch_add(1);
            if (y >= height) break;
            fn(y, per[tid]);
        }
    };
    if (nthreads == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto &t : th) t.join();
}

void write_counters(const std::vector<Counters> &per, uint64_t *counters) {
    if (!counters) return;
    Counters c;
    for (auto &x : per) c.add(x);
    counters[0] += c.closest;
    counters[1] += c.shadow;
    counters[2] += c.discarded;
    counters[3] += c.nodes;
    counters[4] += c.tris;
}

}  // namespace

void orc_render_pt(const orc_scene *s, const orc_camera *cam, const orc_params *p, float *color, float *normal,
                   float *albedo, float *position, int nthreads, uint64_t *counters) {
    std::vector<Counters> per;
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        for (int x = 0; x < p->width; x++) {
            RayPacket rp;
            pt_sample(cx, cam, p, (uint32_t)x, (uint32_t)y, rp);
            size_t i = (size_t)y * p->width + x;
            float col[4] = {rp.value.x, rp.value.y, rp.value.z, 1.0f};
            if (isnan3(rp.value)) {  // amadeus_path_tracer.cu:118-120
                col[0] = 1;
                col[1] = 0;
                col[2] = 1;
                col[3] = 1;
            }
            float nrm[4] = {rp.normal.x, rp.normal.y, rp.normal.z, 0.0f};
            float alb[4] = {rp.albedo.x, rp.albedo.y, rp.albedo.z, 0.0f};
            accumulate4(color + 4 * i, col, p);
            accumulate4(normal + 4 * i, nrm, p);
            accumulate4(albedo + 4 * i, alb, p);
            position[4 * i] = rp.position.x;
            position[4 * i + 1] = rp.position.y;
            position[4 * i + 2] = rp.position.z;
            position[4 * i + 3] = 0.0f;
        }
        c.add(cx.c);
    });
    write_counters(per, counters);
}

void orc_render_pt_pixels(const orc_scene *s, const orc_camera *cam, const orc_params *p, const int32_t *pix, int n,
                          float *out) {
    Ctx cx{s, Counters()};
    for (int i = 0; i < n; i++) {
        RayPacket rp;
        uint32_t x = (uint32_t)(pix[i] % p->width), y = (uint32_t)(pix[i] / p->width);
        pt_sample(cx, cam, p, x, y, rp);
        bool bad = isnan3(rp.value);
        out[4 * i] = bad ? 1.0f : rp.value.x;
        out[4 * i + 1] = bad ? 0.0f : rp.value.y;
        out[4 * i + 2] = bad ? 1.0f : rp.value.z;
        out[4 * i + 3] = 1.0f;
    }
}

void orc_accumulate_pt_f64(const orc_scene *s, const orc_camera *cam, const orc_params *p, int first, int count,
                           double *acc, int nthreads) {
    std::vector<Counters> per;
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        orc_params q = *p;
        for (int f = first; f < first + count; f++) {
            q.samples = f;
            for (int x = 0; x < p->width; x++) {
                RayPacket rp;
                pt_sample(cx, cam, &q, (uint32_t)x, (uint32_t)y, rp);
                size_t i = (size_t)y * p->width + x;
                if (isnan3(rp.value)) continue;
                acc[3 * i] += rp.value.x;
                acc[3 * i + 1] += rp.value.y;
                acc[3 * i + 2] += rp.value.z;
            }
        }
        c.add(cx.c);
    });
}

// ---------------------------------------------------------------------------
// ReSTIR DI — source/optix/amadeus_restir.cu, include/amadeus/reservoir.cuh
// ---------------------------------------------------------------------------
namespace {

struct Sample {  // DirectLightingSample, reservoir.cuh:78-83
    f3 Le{0, 0, 0}, normal{0, 0, 0}, point{0, 0, 0};
    int type = 0;
};

struct Reservoir {  // reservoir.cuh:19-76
    Sample data;
    float w = 0.0f, W = 0.0f;
    int M = 0;
    f3 seed{0, 0, 0};
    bool update(const Sample &s, float weight) {
        w += weight;
        float eta = rand_uniform(seed);
        bool selected = (eta * w) < weight;
        if (selected) data = s;
        M++;
        return selected;
    }
    void resample(float target) {
        float d = target * (float)M;
        W = (d > 0.0f) ? w / d : 0.0f;
    }
    void reset() {
        w = 0.0f;
        M = 0;
        W = 0.0f;
    }
};

// direct_unoccluded / direct_occluded, amadeus_common.cuh:380-421
f3 direct_li(const SurfaceHit &sh, f3 Le, f3 normal, int type, f3 D, float d) {
    f3 rho = brdf(sh, D, eDiffuse);
    float geometric = fabsf(dot(sh.n, D));
    if (type != 2) {
        float ldot = fabsf(dot(normal, D));
        geometric *= ldot / (d * d);
    }
    return rho * Le * geometric;
}

f3 direct_occ(Ctx &cx, const SurfaceHit &sh, f3 Le, f3 normal, int type, f3 D, float d, bool *occ_out = nullptr) {
    cx.c.shadow++;
    bool occ = occluded(cx.s, sh.x, D, d - kShadowEps, cx.c);
    if (occ_out) *occ_out = occ;
    if (occ) return mk(0, 0, 0);
    return direct_li(sh, Le, normal, type, D, d);
}

// target / occluded_target, amadeus_restir.cu:148-178
float target(const Sample &s, const SurfaceHit &hit) {
    f3 D = s.point - hit.x;
    float d = length(D);
    f3 Li = direct_li(hit, s.Le, s.normal, s.type, D / d, d);
    return Li.x + Li.y + Li.z;
}

float occluded_target(Ctx &cx, const Sample &s, const SurfaceHit &hit) {
    f3 D = s.point - hit.x;
    float d = length(D);
    f3 Li = direct_occ(cx, hit, s.Le, s.normal, s.type, D / d, d);
    return Li.x + Li.y + Li.z;
}

struct RestirRP {
    f3 value{0, 0, 0}, position{0, 0, 0}, normal{0, 0, 0}, albedo{0, 0, 0}, seed{0, 0, 0};
    int depth = 0;
};

}  // namespace

struct orc_restir {
    int w, h;
    std::vector<Reservoir> buf[2];
    int cur = 0;  // index of `current`; previous = 1 - cur
    std::vector<Material> gmat;
    std::vector<float> intermediate, auxiliary;  // float4 per pixel
    float prev_pv[16];
    bool have_prev_cam = false;
};

orc_restir *orc_restir_create(int width, int height, uint32_t user_seed) {
    orc_restir *r = new orc_restir();
    r->w = width;
    r->h = height;
    size_t n = (size_t)width * height;
    for (int b = 0; b < 2; b++) {
        r->buf[b].resize(n);
        for (int y = 0; y < height; y++)
            for (int x = 0; x < width; x++) r->buf[b][(size_t)y * width + x].seed = reservoir_seed(x, y, user_seed);
    }
    r->gmat.resize(n);
    for (auto &m : r->gmat) m = Material{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), 0.0f, 1.0f, 0};
    r->intermediate.assign(4 * n, 0.0f);
    r->auxiliary.assign(4 * n, 0.0f);
    std::fill(r->prev_pv, r->prev_pv + 16, 0.0f);
    return r;
}

void orc_restir_destroy(orc_restir *r) { delete r; }

namespace {

// __closesthit__initial (amadeus_restir.cu:437-559) and the recursive bounces.
void restir_trace(Ctx &cx, orc_restir *r, const orc_params *p, int index, f3 o, f3 d, RestirRP &rp) {
    if (rp.depth > p->max_depth) {
        cx.c.discarded++;
        rp.value = mk(0, 0, 0);
        return;
    }
    cx.c.closest++;
    Hit h = closest(cx.s, o, d, kTraceTmax, 2, cx.c);
    if (h.prim < 0) {  // __miss__ (amadeus_restir.cu:566-577): the environment map at (u, v), or 0
        const EnvMap env = cx.s->envmap();
        rp.value = env.tex ? env_radiance(env, d, false) : mk(0, 0, 0);
        return;
    }
    Surface sf;
    if (!load_surface(cx.s, h, d, sf)) {
        rp.value = mk(0, 0, 0);
        return;
    }
    bool primary = rp.depth == 0;
    const Material &material = sf.mat;
    f3 n = sf.n;
    f3 x = sf.x + ((material.type == eTransmission ? -1.0f : 1.0f) * n) * kEps;
    SurfaceHit sh{material, sf.entering, n, -d, x};
    f3 direct = mk(0, 0, 0);
    const orc_scene *s = cx.s;
    if (primary) {
        Reservoir &res = r->buf[r->cur][index];
        res.reset();
        const EnvMap env = s->envmap();
        const unsigned nl = (unsigned)s->lights.size();
        unsigned total = nl + (env.tex ? 1u : 0u);
        for (int i = 0; i < p->candidates; i++) {
            // sample_direct, amadeus_common.cuh:301-377 (triangle lights, the environment)
            Sample smp;
            float pdf = 0.0f;
            if (total > 0) {
                unsigned li = pick_light(rp.seed, total);
                if (li < nl) {
                    const TriLight &L = s->lights[li];
                    smp.point = sample_area_light(L, rp.seed);
                    smp.Le = L.intensity;
                    smp.normal = light_normal(L);
                    smp.type = 1;
                    pdf = 1.0f / (light_area(L) * (float)total);
                } else {  // amadeus_common.cuh:341-370: a point WORLD_RADIUS away
                    f3 wi = env_sample_dir(rp.seed);
                    smp.point = sh.x + wi * kEnvDistance;
                    smp.Le = env_radiance(env, wi, false);
                    smp.normal = -wi;
                    smp.type = 2;
                    pdf = 1.0f / (kFourPiF * (float)total);
                }
            } else {
                rand_uniform(rp.seed);
            }
            f3 D = smp.point - sh.x;
            float dd = length(D);
            f3 Li = direct_li(sh, smp.Le, smp.normal, smp.type, D / dd, dd);
            float tgt = Li.x + Li.y + Li.z;
            res.update(smp, (pdf > 0.0f) ? tgt / pdf : 0.0f);
        }
        Sample smp = res.data;
        f3 D = smp.point - sh.x;
        float dd = length(D);
        bool occ;
        f3 Li = direct_occ(cx, sh, smp.Le, smp.normal, smp.type, D / dd, dd, &occ);
        res.resample(Li.x + Li.y + Li.z);
        // amadeus_restir.cu:522-523 traces the identical shadow ray again
        res.W *= 1.0f - (occ ? 1.0f : 0.0f);
        r->gmat[index] = material;
    } else {
        direct = material.emission + Ld(cx, sh, rp.seed);
    }
    int out;
    f3 wi;
    float pdf;
    f3 f = eval_bsdf(sh, wi, pdf, out, rp.seed);
    f3 T = (pdf > 0.0f) ? f * fabsf(dot(wi, n)) / pdf : mk(0, 0, 0);
    rp.depth++;
    f3 indirect = mk(0, 0, 0);
    if (pdf > 0.0f) {
        restir_trace(cx, r, p, index, x, wi, rp);
        indirect = rp.value;
    }
    rp.value = direct + T * indirect;
    rp.position = x;
    rp.normal = n;
    rp.albedo = material.diffuse;
}

SurfaceHit gbuffer_hit(const orc_restir *r, const orc_camera *cam, const float *normal, const float *position,
                       int index, f3 wo_of) {
    (void)wo_of;
    SurfaceHit hit;
    hit.mat = r->gmat[index];
    hit.entering = false;
    hit.n = load3(normal + 4 * (size_t)index);
    hit.x = load3(position + 4 * (size_t)index);
    hit.wo = normalize(load3(cam->center) - hit.x);
    return hit;
}

// reproject, amadeus_restir.cu:182-210 (ix/iy clamped to the image)
int reproject(const orc_restir *r, const orc_params *p, int index, f3 pos) {
    if (p->samples > 0) return index;
    const float *m = r->prev_pv;
    float v[4];
    for (int k = 0; k < 4; k++)
        v[k] = (m[0 + k] * pos.x + m[4 + k] * pos.y) + (m[8 + k] * pos.z + m[12 + k] * 1.0f);
    float u = v[0] / v[3], vv = v[1] / v[3];
    if (!(u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f)) return -1;
    u = (u + 1.0f) * 0.5f;
    vv = (vv + 1.0f) * 0.5f;
    int ix = (int)((double)(u * (float)p->width) + 0.5);
    int iy = (int)((double)(vv * (float)p->height) + 0.5);
    ix = std::min(std::max(ix, 0), p->width - 1);
    iy = std::min(std::max(iy, 0), p->height - 1);
    return iy * p->width + ix;
}

// sample_spatial_neighborhood, amadeus_restir.cu:295-314
int spatial_neighbour(const orc_params *p, int index, f3 &seed, int radius) {
    int width = p->width, height = p->height;
    int x = index % width, y = index / width;
    f3 off = rand_uniform_3f(seed);
    float c, s;
    sincos_turn(off.x, c, s);
    float rr = (float)radius * sqrtf(off.y);
    int x0 = (int)((float)x + rr * c);
    int y0 = (int)((float)y + rr * s);
    x0 = std::min(std::max(x0, 0), width - 1);
    y0 = std::min(std::max(y0, 0), height - 1);
    return x0 + y0 * width;
}

}  // namespace

void orc_restir_frame(orc_restir *r, const orc_scene *s, const orc_camera *cam, const orc_params *p, float *color,
                      float *normal, float *albedo, float *position, int nthreads, uint64_t *counters) {
    const int W = p->width;
    std::vector<Counters> all, per;
    // Initial pass: __raygen__initial (amadeus_restir.cu:82-144)
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        for (int x = 0; x < W; x++) {
            int index = y * W + x;
            r->buf[r->cur][index].reset();
            RestirRP rp;
            rp.seed = pixel_seed((uint32_t)x, (uint32_t)y, (uint32_t)p->samples, p->user_seed);
            f3 o, d;
            make_ray(cam, p, (uint32_t)x, (uint32_t)y, rp.seed, false, o, d);
            restir_trace(cx, r, p, index, o, d, rp);
            float col[4] = {rp.value.x, rp.value.y, rp.value.z, 1.0f};
            if (isnan3(rp.value)) {
                col[0] = 1;
                col[1] = 0;
                col[2] = 1;
            }
            float *im = &r->intermediate[4 * (size_t)index];
            for (int k = 0; k < 4; k++) im[k] = col[k];
            float *ax = &r->auxiliary[4 * (size_t)index];
            ax[0] = rp.seed.x;
            ax[1] = rp.seed.y;
            ax[2] = rp.seed.z;
            ax[3] = 1.0f;
            float nrm[4] = {rp.normal.x, rp.normal.y, rp.normal.z, 0.0f};
            float alb[4] = {rp.albedo.x, rp.albedo.y, rp.albedo.z, 0.0f};
            accumulate4(normal + 4 * (size_t)index, nrm, p);
            accumulate4(albedo + 4 * (size_t)index, alb, p);
            float *ps = position + 4 * (size_t)index;
            ps[0] = rp.position.x;
            ps[1] = rp.position.y;
            ps[2] = rp.position.z;
            ps[3] = 0.0f;
        }
        c.add(cx.c);
    });
    all.insert(all.end(), per.begin(), per.end());
    // Temporal pass: __raygen__temporal (amadeus_restir.cu:213-291)
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        std::vector<Reservoir> &current = r->buf[r->cur];
        std::vector<Reservoir> &previous = r->buf[1 - r->cur];
        for (int x = 0; x < W; x++) {
            int index = y * W + x;
            Reservoir &cr = current[index];
            if (cr.M <= 0) continue;
            SurfaceHit hit = gbuffer_hit(r, cam, normal, position, index, mk(0, 0, 0));
            int pindex = reproject(r, p, index, hit.x);
            Reservoir *prev = pindex >= 0 ? &previous[pindex] : nullptr;
            Reservoir merged;
            merged.seed = load3(&r->auxiliary[4 * (size_t)index]);
            float t_current = occluded_target(cx, cr.data, hit);
            merged.update(cr.data, t_current * cr.W * (float)cr.M);
            int Nn = cr.M;
            if (prev && prev->M > 0) {
                prev->M = std::min(prev->M, p->m_cap);
                float t_prev = occluded_target(cx, prev->data, hit);
                merged.update(prev->data, t_prev * prev->W * (float)prev->M);
                Nn += (t_prev > 0.0f ? 1 : 0) * prev->M;
            }
            merged.M = Nn;
            merged.resample(target(merged.data, hit));
            cr = merged;
        }
        c.add(cx.c);
    });
    all.insert(all.end(), per.begin(), per.end());
    // Spatial passes: __raygen__spatial (amadeus_restir.cu:317-434) with the
    // host loop and swaps of restir.cuh:282-300.
    for (int pass = 0; pass < p->spatial_passes; pass++) {
        parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
            Ctx cx{s, Counters()};
            const std::vector<Reservoir> &curbuf = r->buf[r->cur];
            std::vector<Reservoir> &outbuf = r->buf[1 - r->cur];
            const int S = p->neighbours;
            for (int x = 0; x < W; x++) {
                int index = y * W + x;
                SurfaceHit hit = gbuffer_hit(r, cam, normal, position, index, mk(0, 0, 0));
                Reservoir current = curbuf[index];
                f3 direct = mk(0, 0, 0);
                if (current.M > 0) {
                    Reservoir merged;
                    merged.seed = load3(&r->auxiliary[4 * (size_t)index]);
                    float t_current = target(current.data, hit);
                    merged.update(current.data, t_current * current.W * (float)current.M);
                    SurfaceHit hits[16];
                    int sizes[16];
                    int count = 0;
                    for (int i = 0; i < S; i++) {
                        int index0 = spatial_neighbour(p, index, merged.seed, p->radius);
                        const Reservoir &nb = curbuf[index0];
                        float t_nb = occluded_target(cx, nb.data, hit);
                        if (nb.M > 0) merged.update(nb.data, t_nb * nb.W * (float)nb.M);
                        hits[i].mat = r->gmat[index0];
                        hits[i].entering = nb.M > 0;
                        hits[i].n = load3(normal + 4 * (size_t)index0);
                        hits[i].wo = hit.wo;
                        hits[i].x = load3(position + 4 * (size_t)index0);
                        sizes[i] = nb.M;
                        count += nb.M;
                    }
                    merged.M = current.M + count;
                    int Z = current.M;
                    for (int i = 0; i < S; i++) {
                        float t_nb = hits[i].entering ? occluded_target(cx, merged.data, hits[i]) : 0.0f;
                        if (t_nb > 0.0f) Z += sizes[i];
                    }
                    float t_merged = target(merged.data, hit);
                    float denominator = t_merged * (float)Z;
                    merged.W = (denominator > 0.0f) ? merged.w / denominator : 0.0f;
                    const Sample &smp = merged.data;
                    f3 D = smp.point - hit.x;
                    float d = length(D);
                    f3 lighting = direct_occ(cx, hit, smp.Le, smp.normal, smp.type, D / d, d);
                    direct = hit.mat.emission + lighting * merged.W;
                    outbuf[index] = merged;
                } else {
                    outbuf[index].reset();
                }
                const float *im = &r->intermediate[4 * (size_t)index];
                float col[4] = {direct.x + im[0], direct.y + im[1], direct.z + im[2], 1.0f + im[3]};
                if (std::isnan(col[0]) || std::isnan(col[1]) || std::isnan(col[2]) || std::isnan(col[3])) {
                    col[0] = 1;
                    col[1] = 0;
                    col[2] = 1;
                    col[3] = 1;
                }
                accumulate4(color + 4 * (size_t)index, col, p);
            }
            c.add(cx.c);
        });
        all.insert(all.end(), per.begin(), per.end());
        r->cur = 1 - r->cur;  // swap(current, previous)
    }
    r->cur = 1 - r->cur;  // final swap (restir.cuh:300)
    std::copy(cam->pv, cam->pv + 16, r->prev_pv);
    r->have_prev_cam = true;
    write_counters(all, counters);
}

void orc_restir_reservoirs(const orc_restir *r, int which, float *out) {
    const std::vector<Reservoir> &b = r->buf[which == 0 ? 1 - r->cur : r->cur];
    for (size_t i = 0; i < b.size(); i++) {
        const Reservoir &q = b[i];
        float *o = out + 16 * i;
        o[0] = q.data.Le.x;
        o[1] = q.data.Le.y;
        o[2] = q.data.Le.z;
        o[3] = q.data.normal.x;
        o[4] = q.data.normal.y;
        o[5] = q.data.normal.z;
        o[6] = q.data.point.x;
        o[7] = q.data.point.y;
        o[8] = q.data.point.z;
        o[9] = (float)q.data.type;
        o[10] = q.w;
        o[11] = q.W;
        o[12] = (float)q.M;
        o[13] = q.seed.x;
        o[14] = q.seed.y;
        o[15] = q.seed.z;
    }
}

// ---------------------------------------------------------------------------
// ReSTIR GI (no reference implementation exists; Ouyang et al. 2021, "ReSTIR
// GI: Path Resampling for Real-Time Path Tracing").  Our algorithm, restated
// identically by the HIP kernels:
//  initial : primary hit x_v (fixed -0.5 pixel offset as the ReSTIR raygen),
//            direct = Le + Ld (path-tracer NEE), one BSDF-sampled direction;
//            its hit x_s with outgoing radiance L_o = path-tracer value of the
//            rest of the path (levels 1..max_depth) is the sample.
//            p_hat = sum(L_o) (zero if x_s is below x_v's tangent plane),
//            w = p_hat / pdf.
//  temporal: merge with the previous frame's reservoir (same pixel while
//            accumulating, reprojected otherwise), M capped, reconnection
//            Jacobian |cos phi_r| / |cos phi_q| * |x_q - x_s|^2 / |x_r - x_s|^2,
//            visibility x_v -> x_s in the target of reused samples.
//  spatial : `neighbours` reservoirs in a `radius` disk with normal/depth
//            similarity, Jacobian and visibility; 1/Z bias correction.
//  shading : direct + f(x_v, wo, wi_s) * L_o * |n_v . wi_s| * W.
// ---------------------------------------------------------------------------
namespace {

struct GISample {
    f3 xs{0, 0, 0}, ns{0, 0, 0}, Lo{0, 0, 0};
};

struct GIRes {
    GISample s;
    f3 xo{0, 0, 0};
    float w = 0.0f, W = 0.0f;
    int M = 0;
    f3 seed{0, 0, 0};
    bool update(const GISample &x, float weight) {
        w += weight;
        float eta = rand_uniform(seed);
        bool sel = (eta * w) < weight;
        if (sel) s = x;
        M++;
        return sel;
    }
};

struct GIG {  // G-buffer
    int valid = 0;
    f3 x{0, 0, 0}, n{0, 0, 0};
    Material mat{};
    float depth = 0.0f;
};

const float kGiMaxJacobian = 10.0f;
const float kGiNormalCos = 0.9f;
const float kGiDepthRel = 0.1f;

float gi_target(const GISample &s, f3 xv, f3 nv) {
    f3 D = s.xs - xv;
    if (!(dot(nv, D) > 0.0f)) return 0.0f;
    return s.Lo.x + s.Lo.y + s.Lo.z;
}

float gi_jacobian(f3 xq, f3 xr, const GISample &s) {
    f3 dq = xq - s.xs, dr = xr - s.xs;
    float lq2 = dot(dq, dq), lr2 = dot(dr, dr);
    if (!(lq2 > 0.0f) || !(lr2 > 0.0f)) return 0.0f;
    float cq = fabsf(dot(s.ns, dq)) / sqrtf(lq2);
    float cr = fabsf(dot(s.ns, dr)) / sqrtf(lr2);
    if (!(cq > 0.0f)) return 0.0f;
    float J = (cr / cq) * (lq2 / lr2);
    return (J > kGiMaxJacobian) ? 0.0f : J;
}

bool gi_visible(Ctx &cx, f3 xv, f3 xs) {
    f3 D = xs - xv;
    float d = length(D);
    cx.c.shadow++;
    return !occluded(cx.s, xv, D / d, d - kShadowEps, cx.c);
}

bool gi_similar(const GIG &a, const GIG &b) {
    return b.valid && dot(a.n, b.n) > kGiNormalCos && fabsf(a.depth - b.depth) < kGiDepthRel * a.depth;
}

}  // namespace

struct orc_gi {
    int w, h;
    std::vector<GIRes> buf[2];
    int cur = 0;
    std::vector<GIG> g[2];
    int gcur = 0;
    std::vector<f3> direct, aux;
    float prev_pv[16];
};

orc_gi *orc_gi_create(int width, int height, uint32_t user_seed) {
    orc_gi *r = new orc_gi();
    r->w = width;
    r->h = height;
    size_t n = (size_t)width * height;
    for (int b = 0; b < 2; b++) {
        r->buf[b].resize(n);
        r->g[b].resize(n);
        for (int y = 0; y < height; y++)
            for (int x = 0; x < width; x++)
                r->buf[b][(size_t)y * width + x].seed = reservoir_seed(x, y, user_seed ^ 0x47490000u);
    }
    r->direct.assign(n, mk(0, 0, 0));
    r->aux.assign(n, mk(0, 0, 0));
    std::fill(r->prev_pv, r->prev_pv + 16, 0.0f);
    return r;
}

void orc_gi_destroy(orc_gi *r) { delete r; }

void orc_gi_frame(orc_gi *r, const orc_scene *s, const orc_camera *cam, const orc_params *p, float *color,
                  int nthreads, uint64_t *counters) {
    const int W = p->width;
    std::vector<Counters> all, per;
    std::vector<GIG> &G = r->g[r->gcur];
    const std::vector<GIG> &PG = r->g[1 - r->gcur];
    f3 center = load3(cam->center);
    // initial
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        for (int x = 0; x < W; x++) {
            int pix = y * W + x;
            GIRes &R = r->buf[r->cur][pix];
            R.w = 0.0f;
            R.M = 0;
            R.W = 0.0f;
            f3 seed = pixel_seed((uint32_t)x, (uint32_t)y, (uint32_t)p->samples, p->user_seed);
            f3 o, d;
            make_ray(cam, p, (uint32_t)x, (uint32_t)y, seed, false, o, d);
            GIG gb;
            r->direct[pix] = mk(0, 0, 0);
            cx.c.closest++;
            Hit h = closest(s, o, d, kTraceTmax, 2, cx.c);
            Surface sf;
            if (h.prim >= 0 && load_surface(s, h, d, sf)) {
                const Material &mat = sf.mat;
                f3 n = sf.n;
                f3 xv = sf.x + ((mat.type == eTransmission ? -1.0f : 1.0f) * n) * kEps;
                SurfaceHit sh{mat, sf.entering, n, -d, xv};
                r->direct[pix] = mat.emission + Ld(cx, sh, seed);
                int out;
                f3 wi;
                float pdf;
                eval_bsdf(sh, wi, pdf, out, seed);
                GISample smp;
                smp.xs = xv + wi;
                smp.ns = -wi;
                if (pdf > 0.0f && p->max_depth >= 1) {
                    RayPacket rp;
                    rp.seed = seed;
                    rp.depth = 1;
                    pt_trace(cx, p->max_depth, xv, wi, rp);
                    seed = rp.seed;
                    if (dot(rp.normal, rp.normal) > 0.0f) {
                        smp.xs = xv + wi * rp.t;  // on the ray, so xs - xv keeps the sampled direction
                        smp.ns = rp.normal;
                        smp.Lo = rp.value;
                    }
                }
                float ph = gi_target(smp, xv, n);
                R.update(smp, (pdf > 0.0f) ? ph / pdf : 0.0f);
                float dd = ph * (float)R.M;
                R.W = (dd > 0.0f) ? R.w / dd : 0.0f;
                R.xo = xv;
                gb.valid = 1;
                gb.x = xv;
                gb.n = n;
                gb.mat = mat;
                gb.depth = length(xv - center);
            }
            G[pix] = gb;
            r->aux[pix] = seed;
        }
        c.add(cx.c);
    });
    all.insert(all.end(), per.begin(), per.end());
    // temporal
    parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        std::vector<GIRes> &current = r->buf[r->cur];
        std::vector<GIRes> &previous = r->buf[1 - r->cur];
        for (int x = 0; x < W; x++) {
            int pix = y * W + x;
            const GIG &gb = G[pix];
            GIRes &R = current[pix];
            if (!gb.valid || R.M <= 0) continue;
            GIRes m;
            m.seed = r->aux[pix];
            m.update(R.s, gi_target(R.s, gb.x, gb.n) * R.W * (float)R.M);
            int N = R.M;
            int q = pix;
            if (p->samples == 0) {
                const float *mm = r->prev_pv;
                float v[4];
                for (int k = 0; k < 4; k++)
                    v[k] = (mm[0 + k] * gb.x.x + mm[4 + k] * gb.x.y) + (mm[8 + k] * gb.x.z + mm[12 + k] * 1.0f);
                float u = v[0] / v[3], vv = v[1] / v[3];
                q = -1;
                if (u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f) {
                    int ix = (int)((double)((u + 1.0f) * 0.5f * (float)p->width) + 0.5);
                    int iy = (int)((double)((vv + 1.0f) * 0.5f * (float)p->height) + 0.5);
                    ix = std::min(std::max(ix, 0), p->width - 1);
                    iy = std::min(std::max(iy, 0), p->height - 1);
                    q = iy * p->width + ix;
                }
            }
            if (q >= 0 && gi_similar(gb, PG[q]) && previous[q].M > 0) {
                GIRes &P = previous[q];
                P.M = std::min(P.M, p->m_cap);
                float J = gi_jacobian(P.xo, gb.x, P.s);
                float t = 0.0f;
                if (J > 0.0f && gi_target(P.s, gb.x, gb.n) > 0.0f && gi_visible(cx, gb.x, P.s.xs))
                    t = gi_target(P.s, gb.x, gb.n);
                m.update(P.s, t * P.W * (float)P.M * J);
                N += (t > 0.0f ? 1 : 0) * P.M;
            }
            m.M = N;
            float tm = gi_target(m.s, gb.x, gb.n) * (float)m.M;
            m.W = (tm > 0.0f) ? m.w / tm : 0.0f;
            m.xo = gb.x;
            R = m;
        }
        c.add(cx.c);
    });
    all.insert(all.end(), per.begin(), per.end());
    // spatial passes (+ shading in the last one; with none, shade the temporal result)
    const int passes = p->spatial_passes;
    auto shade = [&](int pix, const GIRes &R) {
        const GIG &gb = G[pix];
        f3 col = r->direct[pix];
        if (gb.valid && R.W > 0.0f) {
            f3 wi = normalize(R.s.xs - gb.x);
            SurfaceHit sh{gb.mat, true, gb.n, normalize(center - gb.x), gb.x};
            f3 f = brdf(sh, wi, eDiffuse);
            col = col + (f * R.s.Lo * fabsf(dot(gb.n, wi))) * R.W;
        }
        float c4[4] = {col.x, col.y, col.z, 1.0f};
        if (isnan3(col)) {
            c4[0] = 1;
            c4[1] = 0;
            c4[2] = 1;
        }
        accumulate4(color + 4 * (size_t)pix, c4, p);
    };
    if (passes == 0) {
        for (int pix = 0; pix < W * p->height; pix++) shade(pix, r->buf[r->cur][pix]);
    }
    for (int pass = 0; pass < passes; pass++) {
        const bool last = pass == passes - 1;
        parallel_rows(p->height, nthreads, per, [&](int y, Counters &c) {
            Ctx cx{s, Counters()};
            const std::vector<GIRes> &cb = r->buf[r->cur];
            std::vector<GIRes> &ob = r->buf[1 - r->cur];
            for (int x = 0; x < W; x++) {
                int pix = y * W + x;
                const GIG &gb = G[pix];
                GIRes R = cb[pix];
                if (gb.valid && R.M > 0) {
                    GIRes m;
                    m.seed = r->aux[pix];
                    m.update(R.s, gi_target(R.s, gb.x, gb.n) * R.W * (float)R.M);
                    int used[16], Ms[16], nu = 0, cnt = 0;
                    for (int i = 0; i < p->neighbours; i++) {
                        int q = spatial_neighbour(p, pix, m.seed, p->radius);
                        if (q == pix || !gi_similar(gb, G[q])) continue;
                        const GIRes &Nb = cb[q];
                        if (Nb.M <= 0) continue;
                        float J = gi_jacobian(Nb.xo, gb.x, Nb.s);
                        float t = 0.0f;
                        if (J > 0.0f && gi_target(Nb.s, gb.x, gb.n) > 0.0f && gi_visible(cx, gb.x, Nb.s.xs))
                            t = gi_target(Nb.s, gb.x, gb.n);
                        m.update(Nb.s, t * Nb.W * (float)Nb.M * J);
                        used[nu] = q;
                        Ms[nu] = Nb.M;
                        nu++;
                        cnt += Nb.M;
                    }
                    int Z = R.M;
                    for (int k = 0; k < nu; k++) {
                        const GIG &gq = G[used[k]];
                        if (gi_jacobian(gq.x, gb.x, m.s) > 0.0f && gi_target(m.s, gq.x, gq.n) > 0.0f &&
                            gi_visible(cx, gq.x, m.s.xs))
                            Z += Ms[k];
                    }
                    m.M = R.M + cnt;
                    float den = gi_target(m.s, gb.x, gb.n) * (float)Z;
                    m.W = (den > 0.0f) ? m.w / den : 0.0f;
                    m.xo = gb.x;
                    ob[pix] = m;
                    if (last) shade(pix, m);
                } else {
                    GIRes &o = ob[pix];
                    o.w = 0.0f;
                    o.M = 0;
                    o.W = 0.0f;
                    if (last) shade(pix, o);
                }
            }
            c.add(cx.c);
        });
        all.insert(all.end(), per.begin(), per.end());
        r->cur = 1 - r->cur;
    }
    r->cur = 1 - r->cur;  // the last output becomes `previous`
    r->gcur = 1 - r->gcur;
    std::copy(cam->pv, cam->pv + 16, r->prev_pv);
    write_counters(all, counters);
}

void orc_gi_reservoirs(const orc_gi *r, int which, float *out) {
    const std::vector<GIRes> &b = r->buf[which == 0 ? 1 - r->cur : r->cur];
    for (size_t i = 0; i < b.size(); i++) {
        const GIRes &q = b[i];
        float v[20] = {q.s.xs.x, q.s.xs.y, q.s.xs.z, q.s.ns.x, q.s.ns.y, q.s.ns.z, q.s.Lo.x, q.s.Lo.y, q.s.Lo.z,
                       q.xo.x,   q.xo.y,   q.xo.z,   q.w,      q.W,      (float)q.M, q.seed.x, q.seed.y, q.seed.z,
                       0.0f,     0.0f};
        std::copy(v, v + 20, out + 20 * i);
    }
}

float orc_gi_jacobian(const float *xq, const float *xr, const float *xs, const float *ns) {
    GISample s;
    s.xs = load3(xs);
    s.ns = load3(ns);
    return gi_jacobian(load3(xq), load3(xr), s);
}

void orc_camera_rays(const orc_camera *cam, const orc_params *p, int jitter, float *rays) {
    for (int y = 0; y < p->height; y++)
        for (int x = 0; x < p->width; x++) {
            f3 seed = pixel_seed((uint32_t)x, (uint32_t)y, (uint32_t)p->samples, p->user_seed);
            f3 o, d;
            make_ray(cam, p, (uint32_t)x, (uint32_t)y, seed, jitter != 0, o, d);
            float *r = rays + 8 * ((size_t)y * p->width + x);
            r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = kTraceTmax;
            r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 0.0f;
        }
}

// ---------------------------------------------------------------------------
// HDR -> LDR — include/cuda/color.cuh:14-121, include/cuda/cast.cuh:16-20
// ---------------------------------------------------------------------------
// clamp (include/cuda/math.cuh:101-104): fmaxf(a, fminf(f, b)); a NaN input
// therefore clamps to b.
static float ldr_clamp(float f, float a, float b) { return std::fmax(a, std::fmin(f, b)); }

// aces (color.cuh:35-41), per channel: a = c (2.51 c + 0.03),
// b = c (2.43 c + 0.59) + 0.14, clamp(a / b, 0, 1)
static float ldr_aces(float c) {
    float a = c * (2.51f * c + 0.03f);
    float b = c * (2.43f * c + 0.59f) + 0.14f;
    return ldr_clamp(a / b, 0.0f, 1.0f);
}

// srgb (color.cuh:14-31): 12.92 c below 0.0031308, else 1.055 c^(1/2.4) - 0.055
static float ldr_srgb(float c) {
    const float inv_gamma = 1.0f / 2.4f;
    float powed = std::pow(c, inv_gamma);
    float s = c < 0.0031308f ? 12.92f * c : 1.055f * powed - 0.055f;
    return ldr_clamp(s, 0.0f, 1.0f);
}

// quantize_unsigned_8_bits (color.cuh:43-56): min((uint)(x * 256), 255)
static uint32_t ldr_quantize(float x) {
    x = ldr_clamp(x, 0.0f, 1.0f);
    uint32_t q = (uint32_t)(x * 256.0f);
    return q < 255u ? q : 255u;
}

// hdr_to_ldr_kernel (color.cuh:82-100): out row = height - y - 1; to_ui32
// packs x | y << 8 | z << 16 | w << 24 with w = 255 (make_color, :58-67).
void orc_hdr_to_ldr(const float *pixels, uint32_t *target, int width, int height, int tonemapping) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            const float *p = pixels + 4 * ((size_t)y * width + x);
            uint32_t ch[3];
            for (int k = 0; k < 3; k++) ch[k] = ldr_quantize(tonemapping == 0 ? ldr_aces(p[k]) : ldr_srgb(p[k]));
            target[(size_t)(height - y - 1) * width + x] = ch[0] | (ch[1] << 8) | (ch[2] << 16) | (255u << 24);
        }
}

// ---------------------------------------------------------------------------
// Mamba: the editor's G-buffer ReSTIR DI (editor/optix/mamba_shader.cu:35-505,
// host sequence editor/mamba.cu:1055-1150).  A second ReSTIR semantics beside
// amadeus_restir.cu: 8 light candidates per pixel drawn by sample_light
// (:99-162: an area light uniformly, then one of its triangles uniformly, pdf
// 1 / (triangle area x lights), the triangle choice left out as in the
// reference), visibility in the initial resampling target, temporal reuse
// with the previous reservoir capped at M = 200, spatial reuse of 2
// neighbours in a 20 px disk with the 1/Z correction, shading ray epsilon
// 1e-3 (:38-63).  Deviations, each deterministic: the raster G-buffer is a
// ray-cast one (pixel-centre camera rays in the armada's orientation); seeds
// (x, y, time) use time = frame number + 4096 x user seed instead of the wall
// clock; reservoir seeds start from reservoir_seed(); the spatial pass reads
// the reservoirs as the temporal pass left them and writes a second buffer
// (the reference overwrites them in place while neighbours read them), and
// the previous-frame reservoirs are double-buffered for the same reason; the
// in-place M cap of a reprojected reservoir is applied to the reader's copy
// only (every reader caps it to the same value); reprojection indices are
// clamped to the image.
// ---------------------------------------------------------------------------
namespace {

struct MLight {  // LightInfo (mamba_shader.cuh:19-25)
    f3 position{0, 0, 0}, normal{0, 0, 0}, emission{0, 0, 0};
    float area = 0.0f;
    bool sky = false;
};

struct MRes {  // Reservoir<LightInfo> (include/amadeus/reservoir.cuh:19-76)
    MLight data;
    float w = 0.0f, W = 0.0f;
    int M = 0;
    f3 seed{0, 0, 0};
    bool update(const MLight &s, float weight) {
        w += weight;
        float eta = rand_uniform(seed);
        bool selected = (eta * w) < weight;
        if (selected) data = s;
        M++;
        return selected;
    }
    void resample(float target) {
        float d = target * (float)M;
        W = (d > 0.0f) ? w / d : 0.0f;
    }
    void reset() {
        w = 0.0f;
        M = 0;
        W = 0.0f;
    }
};

struct MGb {  // G-buffer texel: position, normal (w: entering), triangle, material (-1: no surface), uv
    f3 x{0, 0, 0}, n{0, 0, 0};
    float nw = 0.0f;
    int prim = -1, mat = -1;
    float u = 0.0f, v = 0.0f;
};

const float kMambaEps = 1e-3f;   // mamba_shader.cu:38
const int kMambaCandidates = 8;  // :248
const int kMambaMCap = 200;      // :352, :457
const int kMambaNeighbours = 2;  // :443
const int kMambaRadius = 20;     // :449

inline float mamba_lum(f3 v) { return v.x + v.y + v.z; }  // luminance (:32-37)
inline f3 cleanse(f3 v) { return isnan3(v) ? mk(0, 0, 0) : v; }

}  // namespace

struct orc_mamba {
    int w = 0, h = 0;
    uint32_t user = 0, frame = 0;
    int samples = 0;
    std::vector<MRes> R, S, P[2];
    int pin = 0;  // P[pin] is the previous frame's
    std::vector<MGb> g;
    std::vector<float> Le;  // 3 per pixel
    // area lights: the emissive triangles grouped by material, in order of
    // their first triangle (an entity submesh per material)
    std::vector<int> first, count, tris;
    std::vector<f3> emission;
    float prev_pv[16];
    float prev_cam[12];
    bool have_prev = false;
};

namespace {

float mamba_sample_light(const orc_mamba *m, const orc_scene *s, MLight &li, f3 &seed) {
    const unsigned total = (unsigned)m->first.size();
    if (total == 0) return -1.0f;
    unsigned light_index = (unsigned)(rand_uniform(seed) * (float)total);
    light_index = light_index < total ? light_index : total - 1;  // rand * total may round up to total
    const unsigned cnt = (unsigned)m->count[light_index];
    unsigned ti = (unsigned)(rand_uniform(seed) * (float)cnt);
    ti = ti < cnt ? ti : cnt - 1;
    f3 v0, v1, v2;
    tri_verts(s, m->tris[m->first[light_index] + ti], v0, v1, v2);
    const f3 gnormal = normalize(cross(v1 - v0, v2 - v0));
    const f3 bary = pcg3f(seed);
    float u = bary.x, v = bary.y;
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    li.sky = false;
    li.position = (v0 * (1.0f - u - v) + v1 * u) + v2 * v;
    li.normal = gnormal;
    li.emission = m->emission[light_index];
    li.area = 0.5f * length(cross(v1 - v0, v2 - v0));
    return 1.0f / (li.area * (float)total);
}

f3 mamba_direct(const SurfaceHit &sh, const MLight &li) {  // direct_unoccluded (:70-83)
    const f3 wi = normalize(li.position - sh.x);
    const float R = length(li.position - sh.x);
    const f3 rho = cleanse(brdf(sh, wi, sh.mat.type));
    const float ldotn = li.sky ? 1.0f : fabsf(dot(li.normal, wi));
    const float ndotwi = fabsf(dot(sh.n, wi));
    const float falloff = li.sky ? 1.0f : 1.0f / (R * R);
    return (((rho * li.emission) * ldotn) * ndotwi) * falloff;
}

bool mamba_occluded(Ctx &cx, f3 origin, f3 target) {  // occluded (:38-63)
    cx.c.shadow++;
    const f3 d = normalize(target - origin);
    return occluded(cx.s, origin, d, length(target - origin) - kMambaEps, cx.c);
}

float mamba_target(const SurfaceHit &sh, const MRes &r) { return mamba_lum(mamba_direct(sh, r.data)); }

float mamba_occluded_target(Ctx &cx, const SurfaceHit &sh, const MRes &r) {
    const f3 Li = mamba_direct(sh, r.data);
    return mamba_lum(Li * (1.0f - (mamba_occluded(cx, sh.x, r.data.position) ? 1.0f : 0.0f)));
}

// load_surface_hit (:165-216) from the ray-cast G-buffer
bool mamba_surface(const orc_mamba *m, const orc_scene *s, const orc_camera *cam, int index, SurfaceHit &sh) {
    const MGb &t = m->g[index];
    if (t.mat < 0) return false;
    const f3 origin = load3(cam->center);
    f3 n = t.n;
    if (dot(t.x - origin, n) > 0.0f) n = -n;
    sh.x = t.x;
    sh.wo = normalize(origin - t.x);
    sh.n = normalize(n);
    sh.entering = t.nw > 0.0f;
    sh.mat = s->mats[t.mat];
    // convert_material (editor/optix/common.cuh:34-47): the diffuse texture only
    if (s->textured && s->mat_tex[5 * (size_t)t.mat] >= 0) {
        float t4[4];
        tex_fetch(s->texture(s->mat_tex[5 * (size_t)t.mat]), t.u, t.v, t4);
        sh.mat.diffuse = mk(t4[0], t4[1], t4[2]);
    }
    const float sign = (sh.mat.type == eTransmission) ? -1.0f : 1.0f;
    sh.x = sh.x + (sign * sh.n) * kMambaEps;
    return true;
}

int mamba_reproject(const orc_mamba *m, bool dirty, int index, f3 pos) {  // reproject (:301-330)
    if (!dirty) return index;
    const float *M = m->prev_pv;
    float v[4];
    for (int k = 0; k < 4; k++) v[k] = (M[0 + k] * pos.x + M[4 + k] * pos.y) + (M[8 + k] * pos.z + M[12 + k] * 1.0f);
    float u = v[0] / v[3], vv = v[1] / v[3];
    if (!(u >= -1.0f && u <= 1.0f && vv >= -1.0f && vv <= 1.0f)) return -1;
    u = (u + 1.0f) * 0.5f;
    vv = (vv + 1.0f) * 0.5f;
    int ix = (int)((double)(u * (float)m->w) + 0.5);
    int iy = (int)((double)(vv * (float)m->h) + 0.5);
    ix = std::min(std::max(ix, 0), m->w - 1);
    iy = std::min(std::max(iy, 0), m->h - 1);
    return iy * m->w + ix;
}

int mamba_neighbour(const orc_mamba *m, int index, f3 &seed) {  // sample_spatial_neighborhood (:411-432)
    const int x = index % m->w, y = index / m->w;
    const f3 off = rand_uniform_3f(seed);
    float c, sn;
    sincos_turn(off.x, c, sn);
    const float rr = (float)kMambaRadius * sqrtf(off.y);
    int x0 = (int)((float)x + rr * c);
    int y0 = (int)((float)y + rr * sn);
    x0 = std::min(std::max(x0, 0), m->w - 1);
    y0 = std::min(std::max(y0, 0), m->h - 1);
    return x0 + y0 * m->w;
}

void mamba_shade(orc_mamba *m, Ctx &cx, const SurfaceHit &sh, const MRes &r, int index) {
    const f3 Ld = mamba_direct(sh, r.data);
    const bool occ = mamba_occluded(cx, sh.x, r.data.position);
    const f3 nd = cleanse((Ld * (1.0f - (occ ? 1.0f : 0.0f))) * r.W);
    const float s = (float)m->samples;
    const f3 prev = load3(&m->Le[3 * (size_t)index]);
    const f3 v = ((prev * s + sh.mat.emission) + nd) / (s + 1.0f);
    m->Le[3 * (size_t)index] = v.x;
    m->Le[3 * (size_t)index + 1] = v.y;
    m->Le[3 * (size_t)index + 2] = v.z;
}

}  // namespace

orc_mamba *orc_mamba_create(const orc_scene *s, int width, int height, uint32_t user_seed) {
    auto *m = new orc_mamba();
    m->w = width;
    m->h = height;
    m->user = user_seed;
    const size_t n = (size_t)width * height;
    m->R.resize(n);
    m->S.resize(n);
    m->P[0].resize(n);
    m->P[1].resize(n);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            const size_t i = (size_t)y * width + x;
            const f3 sd = reservoir_seed((uint32_t)x, (uint32_t)y, user_seed);
            m->R[i].seed = m->S[i].seed = m->P[0][i].seed = m->P[1][i].seed = sd;
        }
    m->g.resize(n);
    m->Le.assign(3 * n, 0.0f);
    std::map<uint32_t, int> of_mat;
    for (size_t t = 0; t < s->tri_mat.size(); t++) {
        const Material &mt = s->mats[s->tri_mat[t]];
        if (!(length(mt.emission) > 0.0f)) continue;  // the armada's emissive test (armada.cu:240)
        auto it = of_mat.find(s->tri_mat[t]);
        if (it == of_mat.end()) {
            of_mat[s->tri_mat[t]] = (int)m->emission.size();
            m->emission.push_back(mt.emission);
            m->count.push_back(0);
        }
        m->count[of_mat[s->tri_mat[t]]]++;
    }
    int run = 0;
    for (int c : m->count) {
        m->first.push_back(run);
        run += c;
    }
    m->tris.assign(run, 0);
    std::vector<int> fill(m->count.size(), 0);
    for (size_t t = 0; t < s->tri_mat.size(); t++) {
        auto it = of_mat.find(s->tri_mat[t]);
        if (it == of_mat.end()) continue;
        m->tris[m->first[it->second] + fill[it->second]++] = (int)t;
    }
    std::fill(m->prev_pv, m->prev_pv + 16, 0.0f);
    std::fill(m->prev_cam, m->prev_cam + 12, 0.0f);
    return m;
}

void orc_mamba_destroy(orc_mamba *m) { delete m; }

int orc_mamba_light_count(const orc_mamba *m) { return (int)m->first.size(); }

void orc_mamba_frame(orc_mamba *m, const orc_scene *s, const orc_camera *cam, int temporal, int spatial,
                     int accumulate, float *color, int nthreads) {
    const int W = m->w, H = m->h;
    float camv[12];
    std::memcpy(camv, cam->center, sizeof(float) * 3);
    std::memcpy(camv + 3, cam->U, sizeof(float) * 3);
    std::memcpy(camv + 6, cam->V, sizeof(float) * 3);
    std::memcpy(camv + 9, cam->W, sizeof(float) * 3);
    const bool dirty = m->have_prev && std::memcmp(camv, m->prev_cam, sizeof(camv)) != 0;
    if (!accumulate || dirty) m->samples = 0;
    const float time = (float)m->frame + 4096.0f * (float)m->user;
    std::vector<Counters> per;
    // G-buffer: the surface of the pixel-centre camera ray
    parallel_rows(H, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        for (int x = 0; x < W; x++) {
            const float dx = 2.0f * (((float)x + 0.5f) / (float)W) - 1.0f;
            const float dy = 2.0f * (((float)y + 0.5f) / (float)H) - 1.0f;
            const f3 o = load3(cam->center);
            const f3 d = normalize((dx * load3(cam->U) + dy * load3(cam->V)) + load3(cam->W));
            cx.c.closest++;
            const Hit h = closest(s, o, d, kTraceTmax, 2, cx.c);
            MGb &t = m->g[(size_t)y * W + x];
            t = MGb();
            Surface sf;
            if (h.prim >= 0 && load_surface(s, h, d, sf, false)) {
                t.x = sf.x;
                t.n = sf.n;
                t.nw = sf.entering ? 1.0f : -1.0f;
                t.prim = h.prim;
                t.mat = (int)s->tri_mat[h.prim];
                t.u = sf.u;
                t.v = sf.v;
            }
        }
        c.add(cx.c);
    });
    const bool end_initial = !temporal && !spatial;
    // initial pass (__raygen__direct_primary, :219-266)
    parallel_rows(H, nthreads, per, [&](int y, Counters &c) {
        Ctx cx{s, Counters()};
        for (int x = 0; x < W; x++) {
            const int index = y * W + x;
            MRes &r = m->R[index];
            r.reset();
            SurfaceHit sh;
            if (!mamba_surface(m, s, cam, index, sh)) continue;
            f3 seed = mk((float)x, (float)y, time);
            for (int i = 0; i < kMambaCandidates; i++) {
                MLight li;
                const float pdf = mamba_sample_light(m, s, li, seed);
                const f3 Ld = mamba_direct(sh, li);
                const float t = Ld.x + Ld.y + Ld.z;
                r.update(li, t / pdf > 0.0f ? t / pdf : 0.0f);
            }
            const f3 Ld = mamba_direct(sh, r.data);
            const bool occ = mamba_occluded(cx, sh.x, r.data.position);
            r.resample(mamba_lum(Ld * (1.0f - (occ ? 1.0f : 0.0f))));
            if (end_initial) mamba_shade(m, cx, sh, r, index);
        }
        c.add(cx.c);
    });
    std::vector<MRes> &Pin = m->P[m->pin], &Pout = m->P[1 - m->pin];
    if (temporal) {  // __raygen__temporal_reuse (:333-394)
        const bool end = !spatial;
        parallel_rows(H, nthreads, per, [&](int y, Counters &c) {
            Ctx cx{s, Counters()};
            for (int x = 0; x < W; x++) {
                const int index = y * W + x;
                SurfaceHit sh;
                if (!mamba_surface(m, s, cam, index, sh)) {
                    if (end) {
                        Pout[index] = Pin[index];
                        Pout[index].reset();
                    }
                    continue;
                }
                const int pindex = mamba_reproject(m, dirty, index, sh.x);
                MRes &cur = m->R[index];
                if (cur.M <= 0) {
                    if (end) Pout[index] = Pin[index];
                    continue;
                }
                MRes merged;
                merged.seed = 2.0f * mk((float)x, (float)y, time);
                const float t_cur = mamba_occluded_target(cx, sh, cur);
                merged.update(cur.data, (t_cur * cur.W) * (float)cur.M);
                int N = cur.M;
                if (pindex >= 0 && Pin[pindex].M > 0) {
                    MRes prev = Pin[pindex];
                    prev.M = std::min(prev.M, kMambaMCap);
                    const float t_prev = mamba_occluded_target(cx, sh, prev);
                    merged.update(prev.data, (t_prev * prev.W) * (float)prev.M);
                    N += (t_prev > 0.0f ? 1 : 0) * prev.M;
                }
                merged.M = N;
                merged.resample(mamba_target(sh, merged));
                cur = merged;
                if (end) {
                    mamba_shade(m, cx, sh, cur, index);
                    Pout[index] = cur;
                }
            }
            c.add(cx.c);
        });
    }
    if (spatial) {  // __raygen__spatial_reuse (:435-505)
        parallel_rows(H, nthreads, per, [&](int y, Counters &c) {
            Ctx cx{s, Counters()};
            for (int x = 0; x < W; x++) {
                const int index = y * W + x;
                SurfaceHit sh;
                if (!mamba_surface(m, s, cam, index, sh)) {
                    Pout[index] = Pin[index];
                    Pout[index].reset();
                    m->S[index] = m->R[index];
                    continue;
                }
                const MRes &cur = m->R[index];
                if (cur.M <= 0) {  // the reference's debug colour
                    m->Le[3 * (size_t)index] = 0.0f;
                    m->Le[3 * (size_t)index + 1] = 0.0f;
                    m->Le[3 * (size_t)index + 2] = 1.0f;
                    Pout[index] = Pin[index];
                    Pout[index].reset();
                    m->S[index] = cur;
                    continue;
                }
                MRes merged;
                merged.seed = 2.0f * mk((float)x, (float)y, time);
                const float t_cur = mamba_target(sh, cur);
                merged.update(cur.data, (t_cur * cur.W) * (float)cur.M);
                SurfaceHit hits[kMambaNeighbours];
                int sizes[kMambaNeighbours];
                int count = 0;
                for (int i = 0; i < kMambaNeighbours; i++) {
                    const int index0 = mamba_neighbour(m, index, merged.seed);
                    MRes nb = m->R[index0];
                    nb.M = std::min(nb.M, kMambaMCap);
                    const float t_nb = mamba_occluded_target(cx, sh, nb);
                    if (nb.M > 0) merged.update(nb.data, (t_nb * nb.W) * (float)nb.M);
                    hits[i] = SurfaceHit();
                    mamba_surface(m, s, cam, index0, hits[i]);
                    hits[i].wo = sh.wo;
                    hits[i].entering = nb.M > 0;
                    sizes[i] = nb.M;
                    count += nb.M;
                }
                merged.M = cur.M + count;
                int Z = cur.M;
                for (int i = 0; i < kMambaNeighbours; i++) {
                    const float t = hits[i].entering ? mamba_occluded_target(cx, hits[i], merged) : 0.0f;
                    if (t > 0.0f) Z += sizes[i];
                }
                const float den = mamba_target(sh, merged) * (float)Z;
                merged.W = den > 0.0f ? merged.w / den : 0.0f;
                m->S[index] = merged;
                mamba_shade(m, cx, sh, merged, index);
                Pout[index] = merged;
            }
            c.add(cx.c);
        });
        std::swap(m->R, m->S);
    }
    if (temporal || spatial) m->pin = 1 - m->pin;
    for (size_t i = 0; i < (size_t)W * H; i++) {
        color[4 * i] = m->Le[3 * i];
        color[4 * i + 1] = m->Le[3 * i + 1];
        color[4 * i + 2] = m->Le[3 * i + 2];
        color[4 * i + 3] = 1.0f;
    }
    std::memcpy(m->prev_pv, cam->pv, sizeof(m->prev_pv));
    std::memcpy(m->prev_cam, camv, sizeof(camv));
    m->have_prev = true;
    m->samples++;
    m->frame++;
}

void orc_mamba_reservoirs(const orc_mamba *m, int which, float *out) {
    const std::vector<MRes> &b = which == 0 ? m->R : m->P[m->pin];
    for (size_t i = 0; i < b.size(); i++) {
        const MRes &r = b[i];
        float *o = out + 16 * i;
        o[0] = r.data.position.x, o[1] = r.data.position.y, o[2] = r.data.position.z;
        o[3] = r.data.normal.x, o[4] = r.data.normal.y, o[5] = r.data.normal.z;
        o[6] = r.data.emission.x, o[7] = r.data.emission.y, o[8] = r.data.emission.z;
        o[9] = r.data.area;
        o[10] = r.w, o[11] = r.W, o[12] = (float)r.M;
        o[13] = r.seed.x, o[14] = r.seed.y, o[15] = r.seed.z;
    }
}

// C++ drop-in layer over the krt C ABI: the reference's ArmadaRTX,
// AttachmentRTX, Accelerator, Camera and Transform
// (include/amadeus/armada.cuh:43-336, include/amadeus/accelerator.cuh:133-277,
// include/camera.hpp:15-63, include/transform.hpp) with the same method names
// and argument meaning.  OptiX / Vulkan / glm types become PODs:
// OptixTraversableHandle -> krt_handle, vk::Extent2D -> Extent2D,
// glm::vec3 -> Vec3, glm::mat4 -> Mat4 (column-major, as glm stores it).
// Errors throw std::runtime_error, as the reference's CUDA_CHECK /
// OPTIX_CHECK do (include/cuda/error.cuh:43-98).
//
// AttachmentRTX is the reference's virtual plugin interface (attach, load,
// unload, set_option, get_option, render).  The built-in attachments
// (path_tracer, restir, restir_gi) are subclasses whose render runs the HIP
// passes through krt_armada_render; a user subclass is called by
// ArmadaRTX::render exactly as armada.cu:733-780 calls an attachment, with
// the frame's ArmadaLaunchInfo (camera frame, AOV device buffers, sample
// counter) and the scene handle to trace against.
#ifndef KRT_ARMADA_HPP_
#define KRT_ARMADA_HPP_

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "krt.h"

namespace krt_cpp {

inline void check(int code) {
    if (code != KRT_OK) throw std::runtime_error(std::string("krt: ") + krt_last_error());
}

struct Extent2D {
    uint32_t width, height;
};

// ------------------------------------------------------------------- math --
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 normalize(Vec3 a) { return a * (1.0f / length(a)); }

// glm::mat4 layout: m[column][row]
struct Mat4 {
    float m[4][4] = {};
};
inline Mat4 operator*(const Mat4 &a, const Mat4 &b) {
    Mat4 r;
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += a.m[k][row] * b.m[c][k];
            r.m[c][row] = s;
        }
    return r;
}

// glm::quat(vec3 euler) (w, x, y, z) and its rotation of a vector
struct Quat {
    float w, x, y, z;
};
inline Quat quat_from_euler(Vec3 e) {
    const float cx = std::cos(e.x * 0.5f), cy = std::cos(e.y * 0.5f), cz = std::cos(e.z * 0.5f);
    const float sx = std::sin(e.x * 0.5f), sy = std::sin(e.y * 0.5f), sz = std::sin(e.z * 0.5f);
    return {cx * cy * cz + sx * sy * sz, sx * cy * cz - cx * sy * sz, cx * sy * cz + sx * cy * sz,
            cx * cy * sz - sx * sy * cz};
}
inline Vec3 rotate(const Quat &q, Vec3 v) {  // mat3_cast(q) * v
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z, wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    return {(1.0f - 2.0f * (yy + zz)) * v.x + 2.0f * (xy - wz) * v.y + 2.0f * (xz + wy) * v.z,
            2.0f * (xy + wz) * v.x + (1.0f - 2.0f * (xx + zz)) * v.y + 2.0f * (yz - wx) * v.z,
            2.0f * (xz - wy) * v.x + 2.0f * (yz + wx) * v.y + (1.0f - 2.0f * (xx + yy)) * v.z};
}

// Transform (include/transform.hpp): position, rotation in degrees, scale.
// forward() / up() / right() pass the rotation's numbers to glm::quat
// unconverted, as the reference does (transform.hpp:96-111).
struct Transform {
    Vec3 position, rotation, scale{1.0f, 1.0f, 1.0f};
    Vec3 forward() const { return normalize(rotate(quat_from_euler(rotation), {0.0f, 0.0f, -1.0f})); }
    Vec3 right() const { return normalize(rotate(quat_from_euler(rotation), {1.0f, 0.0f, 0.0f})); }
    Vec3 up() const { return normalize(rotate(quat_from_euler(rotation), {0.0f, 1.0f, 0.0f})); }
};

// Camera (include/camera.hpp:15-38)
struct Camera {
    float fov = 45.0f, aspect = 1.0f;
    explicit Camera(float fov_ = 45.0f, float aspect_ = 1.0f) : fov(fov_), aspect(aspect_) {}
    // glm::perspective(radians(fov), aspect, 0.1, 1000), clip z in [-1, 1]
    Mat4 perspective_matrix() const {
        const float t = std::tan(fov * 0.01745329251994329577f / 2.0f), n = 0.1f, f = 1000.0f;
        Mat4 p;
        p.m[0][0] = 1.0f / (aspect * t);
        p.m[1][1] = 1.0f / t;
        p.m[2][2] = -(f + n) / (f - n);
        p.m[2][3] = -1.0f;
        p.m[3][2] = -(2.0f * f * n) / (f - n);
        return p;
    }
    // glm::lookAt(position, position + forward, up) (right-handed)
    static Mat4 view_matrix(const Transform &t) {
        const Vec3 eye = t.position, f = normalize(t.forward()), s = normalize(cross(f, t.up())), u = cross(s, f);
        Mat4 v;
        v.m[0][0] = s.x, v.m[1][0] = s.y, v.m[2][0] = s.z;
        v.m[0][1] = u.x, v.m[1][1] = u.y, v.m[2][1] = u.z;
        v.m[0][2] = -f.x, v.m[1][2] = -f.y, v.m[2][2] = -f.z;
        v.m[3][0] = -dot(s, eye), v.m[3][1] = -dot(u, eye), v.m[3][2] = dot(f, eye);
        v.m[3][3] = 1.0f;
        return v;
    }
};

// uvw_frame (include/camera.hpp:45-63)
struct UVW {
    Vec3 u, v, w;
};
inline UVW uvw_frame(const Camera &camera, const Transform &transform, float aspect = -1.0f) {
    const Vec3 eye = transform.position, lookat = eye + transform.forward(), up = transform.up();
    const Vec3 w = lookat - eye;
    const float wlen = length(w);
    Vec3 u = normalize(cross(w, up));
    Vec3 v = normalize(cross(u, w));
    const float vlen = wlen * std::tan(camera.fov * 0.01745329251994329577f / 2.0f);
    v = v * vlen;
    const float ulen = vlen * (aspect > 0.0f ? aspect : camera.aspect);
    u = u * ulen;
    return {u, v, w};
}

// ArmadaLaunchInfo (armada.cuh:43-83): what a render call hands its
// attachment.  Lights, materials and the acceleration structure stay behind
// the scene handle; `context` is the krt context to launch on.
struct ArmadaLaunchInfo {
    Extent2D resolution;
    struct {
        Vec3 center, ax_u, ax_v, ax_w;
        Mat4 projection, view;
    } camera;
    struct {
        void *color, *normal, *albedo, *position;  // width x height float4 each, device memory
    } buffers;
    bool accumulate = false;
    int max_depth = 2;
    int samples = 0;
    krt_context *context = nullptr;
};

// The launch's camera as the C ABI takes it: center, U, V, W and
// projection * view (column-major) for reprojection.
inline krt_camera make_camera(const Camera &camera, const Transform &transform, float aspect = -1.0f) {
    const UVW f = uvw_frame(camera, transform, aspect);
    Camera c = camera;
    if (aspect > 0.0f) c.aspect = aspect;
    const Mat4 pv = c.perspective_matrix() * Camera::view_matrix(transform);
    krt_camera k{};
    const Vec3 vs[4] = {transform.position, f.u, f.v, f.w};
    float *dst[4] = {k.center, k.U, k.V, k.W};
    for (int i = 0; i < 4; i++) dst[i][0] = vs[i].x, dst[i][1] = vs[i].y, dst[i][2] = vs[i].z;
    std::memcpy(k.pv, pv.m, sizeof(k.pv));
    return k;
}

class Context {
    krt_context *m_ctx = nullptr;

public:
    explicit Context(int device = 0) { check(krt_create(device, &m_ctx)); }
    ~Context() { krt_destroy(m_ctx); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    krt_context *get() const { return m_ctx; }
};

// amadeus::Accelerator: owns the resident scenes and builds their BVHs on the GPU.
class Accelerator {
    Context &m_ctx;
    std::vector<krt_handle> m_scenes;

public:
    explicit Accelerator(Context &ctx) : m_ctx(ctx) {}
    ~Accelerator() {
        for (krt_handle h : m_scenes) krt_scene_free(m_ctx.get(), h);
    }
    Accelerator(const Accelerator &) = delete;
    Accelerator &operator=(const Accelerator &) = delete;
    // Accelerator::update(entities) + build_tlas(): upload and build; returns
    // the traversable handle.
    krt_handle update(const krt_scene_desc &scene, int leaf_max = 4) {
        krt_handle h = 0;
        check(krt_scene_upload(m_ctx.get(), &scene, &h));
        m_scenes.push_back(h);
        check(krt_accel_build(m_ctx.get(), h, leaf_max));
        return h;
    }
    // The same with material textures and normal maps (convert_material's
    // textures, armada.cu:333-398; Vertex::tex_coords / tangent / bitangent).
    krt_handle update(const krt_scene_desc &scene, const krt_scene_textures &textures, int leaf_max = 4) {
        krt_handle h = 0;
        check(krt_scene_upload_textured(m_ctx.get(), &scene, &textures, &h));
        m_scenes.push_back(h);
        check(krt_accel_build(m_ctx.get(), h, leaf_max));
        return h;
    }
    // A moved entity: new world-space vertices for [first, first + count),
    // then the trees refit to them (same shapes) or rebuilt (after a large
    // motion, where refit boxes grow loose).
    void update_vertices(krt_handle scene, uint32_t first, uint32_t count, const float *positions,
                         const float *normals = nullptr) {
        check(krt_scene_update_vertices(m_ctx.get(), scene, first, count, positions, normals));
    }
    void refit(krt_handle scene) { check(krt_accel_refit(m_ctx.get(), scene)); }
    void rebuild(krt_handle scene, int leaf_max = 4) { check(krt_accel_build(m_ctx.get(), scene, leaf_max)); }
    krt_context *context() const { return m_ctx.get(); }
    Context &ctx() const { return m_ctx; }
};

class ArmadaRTX;

// AttachmentRTX (armada.cuh:107-139)
class AttachmentRTX {
    int m_hit_group_count;

public:
    using OptionValue = std::variant<int, float, bool, std::string>;
    explicit AttachmentRTX(int hit_group_count) : m_hit_group_count(hit_group_count) {}
    virtual ~AttachmentRTX() = default;
    AttachmentRTX(const AttachmentRTX &) = delete;
    AttachmentRTX &operator=(const AttachmentRTX &) = delete;

    virtual void attach(const ArmadaRTX &) = 0;
    virtual void load() = 0;
    virtual void unload() = 0;
    virtual void set_option(const std::string &, const OptionValue &) {}
    virtual OptionValue get_option(const std::string &) const { return {}; }
    virtual void render(const ArmadaRTX *, const ArmadaLaunchInfo &, const std::optional<krt_handle> &,
                        std::vector<krt_handle> *, const Extent2D &) = 0;
    int hit_group_count() const { return m_hit_group_count; }

    friend class ArmadaRTX;
};

// A built-in attachment: the HIP passes behind krt_armada_render.  Its
// load / unload happen inside the armada's render call on an attachment
// switch (armada.cu:747-753), so the overrides only record the state.
class BuiltinAttachment : public AttachmentRTX {
    krt_attachment *m_att = nullptr;
    std::string m_kind;
    krt_armada *m_armada = nullptr;  // set by attach()
    std::map<std::string, OptionValue> m_pending;

public:
    BuiltinAttachment(Context &ctx, const std::string &kind) : AttachmentRTX(1), m_kind(kind) {
        check(krt_attachment_create(ctx.get(), kind.c_str(), &m_att));
    }
    ~BuiltinAttachment() override { krt_attachment_destroy(m_att); }
    krt_attachment *get() const { return m_att; }
    const std::string &kind() const { return m_kind; }
    void attach(const ArmadaRTX &armada) override;
    void load() override {}
    void unload() override {}
    void set_option(const std::string &field, const OptionValue &value) override;
    OptionValue get_option(const std::string &field) const override;
    void render(const ArmadaRTX *, const ArmadaLaunchInfo &, const std::optional<krt_handle> &,
                std::vector<krt_handle> *, const Extent2D &) override {}
};

class PathTracer : public BuiltinAttachment {
public:
    explicit PathTracer(Context &ctx) : BuiltinAttachment(ctx, "path_tracer") {}
};

class ReSTIR : public BuiltinAttachment {
public:
    explicit ReSTIR(Context &ctx) : BuiltinAttachment(ctx, "restir") {}
};

class ReSTIRGI : public BuiltinAttachment {
public:
    explicit ReSTIRGI(Context &ctx) : BuiltinAttachment(ctx, "restir_gi") {}
};

inline krt_option to_option(const AttachmentRTX::OptionValue &value) {
    krt_option o{};
    if (auto p = std::get_if<int>(&value)) o.kind = KRT_OPT_INT, o.i = *p;
    else if (auto q = std::get_if<float>(&value)) o.kind = KRT_OPT_FLOAT, o.f = *q;
    else if (auto b = std::get_if<bool>(&value)) o.kind = KRT_OPT_BOOL, o.b = *b;
    else o.kind = KRT_OPT_STRING, o.s = std::get<std::string>(value).c_str();
    return o;
}

class ArmadaRTX {
    Context &m_ctx;
    krt_armada *m_armada = nullptr;
    Extent2D m_extent;
    std::map<std::string, std::shared_ptr<AttachmentRTX>> m_attachments;
    std::string m_active, m_previous;
    int m_depth = 2;
    int m_samples = 0;            // sample counter of user attachments (built-ins keep theirs in the C layer)
    krt_handle m_last_scene = 0;

public:
    // ArmadaRTX(context, accelerator, mesh daemon, extent) (armada.cu:28-83)
    ArmadaRTX(Accelerator &system, const Extent2D &extent) : m_ctx(system.ctx()), m_extent(extent) {
        check(krt_armada_create(m_ctx.get(), extent.width, extent.height, &m_armada));
    }
    ~ArmadaRTX() { krt_armada_destroy(m_armada); }
    ArmadaRTX(const ArmadaRTX &) = delete;
    ArmadaRTX &operator=(const ArmadaRTX &) = delete;

    krt_armada *get() const { return m_armada; }
    krt_context *context() const { return m_ctx.get(); }
    size_t size() const { return (size_t)m_extent.width * m_extent.height; }
    Extent2D extent() const { return m_extent; }
    std::vector<std::string> attachments() const {
        std::vector<std::string> names;
        for (const auto &kv : m_attachments) names.push_back(kv.first);
        return names;
    }
    std::string active_attachment() const { return m_active; }

    // set_option / get_option forward to the active attachment (armada.cuh:311-320)
    void set_option(const std::string &field, const AttachmentRTX::OptionValue &value) {
        if (m_active.empty()) throw std::runtime_error("krt: no active attachment");
        m_attachments.at(m_active)->set_option(field, value);
    }
    AttachmentRTX::OptionValue get_option(const std::string &field) const {
        if (m_active.empty()) throw std::runtime_error("krt: no active attachment");
        return m_attachments.at(m_active)->get_option(field);
    }

    // device pointers of the float4 AOVs (armada.cuh:287-301)
    krt_buffers buffers() const {
        krt_buffers b{};
        check(krt_armada_buffers(m_armada, &b));
        return b;
    }
    void *color_buffer() const { return buffers().color; }
    void *normal_buffer() const { return buffers().normal; }
    void *albedo_buffer() const { return buffers().albedo; }
    void *position_buffer() const { return buffers().position; }
    // host copy of one AOV (0 color, 1 normal, 2 albedo, 3 position)
    std::vector<float> download(int which) const {
        std::vector<float> out(size() * 4);
        check(krt_armada_readback(m_armada, which, out.data()));
        return out;
    }

    // attach (armada.cuh:303-309): registers and activates the attachment
    void attach(const std::string &name, const std::shared_ptr<AttachmentRTX> &attachment) {
        m_attachments[name] = attachment;
        m_active = name;
        attachment->attach(*this);
        if (auto *b = dynamic_cast<BuiltinAttachment *>(attachment.get())) {
            check(krt_armada_attach(m_armada, name.c_str(), b->get()));
        }
    }
    void activate(const std::string &name) {
        if (!m_attachments.count(name)) throw std::runtime_error("krt: no attachment named " + name);
        m_active = name;
        if (dynamic_cast<BuiltinAttachment *>(m_attachments[name].get()))
            check(krt_armada_activate(m_armada, name.c_str()));
    }
    // ArmadaRTX::set_envmap / set_envmap_enabled (armada.cu:86-96,
    // armada.cuh:323-328), from texels instead of an image path
    void set_envmap(const float *rgba, uint32_t width, uint32_t height) {
        check(krt_armada_set_envmap(m_armada, rgba, width, height));
    }
    void set_envmap_enabled(bool enabled) { check(krt_armada_set_envmap_enabled(m_armada, enabled ? 1 : 0)); }
    void set_depth(int depth) {
        check(krt_armada_set_depth(m_armada, depth));
        m_depth = depth;
    }

    // ArmadaRTX::render(system, entities, transforms, camera, transform,
    // accumulate) (armada.cu:733-780): the camera entity's Camera and
    // Transform give the launch's U, V, W frame (uvw_frame with the camera's
    // aspect, armada.cu:415) and projection * view; the active attachment
    // renders one frame.
    void render(krt_handle scene, const Camera &camera, const Transform &transform, bool accumulate = true) {
        if (m_active.empty()) throw std::runtime_error("krt: no active attachment");
        AttachmentRTX *att = m_attachments.at(m_active).get();
        if (m_active != m_previous) {  // armada.cu:747-753 (a built-in's own load is in krt_armada_render)
            if (!m_previous.empty() && m_attachments.count(m_previous)) m_attachments[m_previous]->unload();
            att->load();
            m_previous = m_active;
        }
        if (dynamic_cast<BuiltinAttachment *>(att)) {
            const krt_camera cam = make_camera(camera, transform);
            check(krt_armada_render(m_armada, scene, &cam, accumulate ? 1 : 0));
            return;
        }
        if (!accumulate || scene != m_last_scene) m_samples = 0;
        m_last_scene = scene;
        ArmadaLaunchInfo info;
        info.resolution = m_extent;
        const UVW f = uvw_frame(camera, transform);
        info.camera.center = transform.position;
        info.camera.ax_u = f.u;
        info.camera.ax_v = f.v;
        info.camera.ax_w = f.w;
        info.camera.projection = camera.perspective_matrix();
        info.camera.view = Camera::view_matrix(transform);
        const krt_buffers b = buffers();
        info.buffers = {b.color, b.normal, b.albedo, b.position};
        info.accumulate = accumulate;
        info.max_depth = m_depth;
        info.samples = m_samples;
        info.context = m_ctx.get();
        att->render(this, info, std::optional<krt_handle>(scene), nullptr, m_extent);
        m_samples++;
    }
    // The same with a precomputed camera frame.
    void render(krt_handle scene, const krt_camera &camera, bool accumulate = true) {
        check(krt_armada_render(m_armada, scene, &camera, accumulate ? 1 : 0));
    }
};

inline void BuiltinAttachment::attach(const ArmadaRTX &armada) { m_armada = armada.get(); }

inline void BuiltinAttachment::set_option(const std::string &field, const OptionValue &value) {
    if (!m_armada) throw std::runtime_error("krt: attach the attachment before setting options");
    const krt_option o = to_option(value);
    check(krt_armada_set_option(m_armada, field.c_str(), &o));
}

inline AttachmentRTX::OptionValue BuiltinAttachment::get_option(const std::string &field) const {
    if (!m_armada) return {};
    krt_option o{};
    check(krt_armada_get_option(m_armada, field.c_str(), &o));
    if (o.kind == KRT_OPT_FLOAT) return o.f;
    if (o.kind == KRT_OPT_BOOL) return (bool)o.b;
    return o.i;
}

// The plugin entry of source/amadeus/naive_path_tracer.cu:280-344
// (`extern "C" ret load_attachment()`), by kind.
struct AttachmentEntry {
    std::string name;
    std::shared_ptr<AttachmentRTX> ptr;
};
inline AttachmentEntry load_attachment(Context &ctx, const std::string &kind) {
    if (kind == "path_tracer") return {"Path Tracer", std::make_shared<PathTracer>(ctx)};
    if (kind == "restir") return {"ReSTIR", std::make_shared<ReSTIR>(ctx)};
    if (kind == "restir_gi") return {"ReSTIR GI", std::make_shared<ReSTIRGI>(ctx)};
    throw std::runtime_error("krt: unknown attachment kind " + kind);
}

// kobra::cuda::hdr_to_ldr (include/cuda/color.cuh:104-118): same arguments,
// plus the context whose stream the launch is queued on.
enum { eTonemappingACES = KRT_TONEMAP_ACES, eTonemappingSRGB = KRT_TONEMAP_SRGB };
inline void hdr_to_ldr(Context &ctx, const void *pixels, uint32_t *target, int width, int height,
                       int tonemapping = eTonemappingACES) {
    check(krt_hdr_to_ldr(ctx.get(), pixels, target, (uint32_t)width, (uint32_t)height, tonemapping));
}

}  // namespace krt_cpp

#endif

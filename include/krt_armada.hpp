// C++ drop-in layer over the krt C ABI: the reference's ArmadaRTX,
// AttachmentRTX and Accelerator classes (include/amadeus/armada.cuh:107-336,
// include/amadeus/accelerator.cuh:133-277) with the same method names and
// argument meaning.  OptiX / Vulkan types become PODs: OptixTraversableHandle
// -> krt_handle, vk::Extent2D -> Extent2D.  Errors throw std::runtime_error,
// as the reference's CUDA_CHECK / OPTIX_CHECK do (include/cuda/error.cuh:43-98).
#ifndef KRT_ARMADA_HPP_
#define KRT_ARMADA_HPP_

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
    // Accelerator::update(entities) + build_tlas(): upload and build; returns
    // the traversable handle.
    krt_handle update(const krt_scene_desc &scene, int leaf_max = 4) {
        krt_handle h = 0;
        check(krt_scene_upload(m_ctx.get(), &scene, &h));
        check(krt_accel_build(m_ctx.get(), h, leaf_max));
        m_scenes.push_back(h);
        return h;
    }
    krt_context *context() const { return m_ctx.get(); }
};

class ArmadaRTX;

// AttachmentRTX (armada.cuh:107-139).  Built-in kinds: "path_tracer", "restir".
class AttachmentRTX {
    krt_attachment *m_att = nullptr;
    std::string m_kind;

public:
    using OptionValue = std::variant<int, float, bool, std::string>;
    AttachmentRTX(Context &ctx, const std::string &kind) : m_kind(kind) {
        check(krt_attachment_create(ctx.get(), kind.c_str(), &m_att));
    }
    virtual ~AttachmentRTX() { krt_attachment_destroy(m_att); }
    AttachmentRTX(const AttachmentRTX &) = delete;
    AttachmentRTX &operator=(const AttachmentRTX &) = delete;
    krt_attachment *get() const { return m_att; }
    const std::string &kind() const { return m_kind; }
};

class PathTracer : public AttachmentRTX {
public:
    explicit PathTracer(Context &ctx) : AttachmentRTX(ctx, "path_tracer") {}
};

class ReSTIR : public AttachmentRTX {
public:
    explicit ReSTIR(Context &ctx) : AttachmentRTX(ctx, "restir") {}
};

struct Camera {  // ArmadaLaunchInfo::Camera (armada.cuh:46-55)
    krt_camera raw;
};

class ArmadaRTX {
    krt_armada *m_armada = nullptr;
    Extent2D m_extent;
    std::map<std::string, std::shared_ptr<AttachmentRTX>> m_attachments;
    std::string m_active;

public:
    ArmadaRTX(Accelerator &system, const Extent2D &extent) : m_extent(extent) {
        check(krt_armada_create(system.context(), extent.width, extent.height, &m_armada));
    }
    ~ArmadaRTX() { krt_armada_destroy(m_armada); }
    ArmadaRTX(const ArmadaRTX &) = delete;
    ArmadaRTX &operator=(const ArmadaRTX &) = delete;

    size_t size() const { return (size_t)m_extent.width * m_extent.height; }
    Extent2D extent() const { return m_extent; }
    std::vector<std::string> attachments() const {
        std::vector<std::string> names;
        for (const auto &kv : m_attachments) names.push_back(kv.first);
        return names;
    }
    std::string active_attachment() const { return m_active; }

    void set_option(const std::string &field, const AttachmentRTX::OptionValue &value) {
        krt_option o{};
        if (auto p = std::get_if<int>(&value)) o.kind = KRT_OPT_INT, o.i = *p;
        else if (auto q = std::get_if<float>(&value)) o.kind = KRT_OPT_FLOAT, o.f = *q;
        else if (auto b = std::get_if<bool>(&value)) o.kind = KRT_OPT_BOOL, o.b = *b;
        else o.kind = KRT_OPT_STRING, o.s = std::get<std::string>(value).c_str();
        check(krt_armada_set_option(m_armada, field.c_str(), &o));
    }
    AttachmentRTX::OptionValue get_option(const std::string &field) const {
        krt_option o{};
        check(krt_armada_get_option(m_armada, field.c_str(), &o));
        if (o.kind == KRT_OPT_FLOAT) return o.f;
        if (o.kind == KRT_OPT_BOOL) return (bool)o.b;
        return o.i;
    }

    // device pointers of the float4 AOVs (armada.cuh:287-301)
    void *color_buffer() { return buffers().color; }
    void *normal_buffer() { return buffers().normal; }
    void *albedo_buffer() { return buffers().albedo; }
    void *position_buffer() { return buffers().position; }
    krt_buffers buffers() {
        krt_buffers b{};
        check(krt_armada_buffers(m_armada, &b));
        return b;
    }

    void attach(const std::string &name, const std::shared_ptr<AttachmentRTX> &attachment) {
        check(krt_armada_attach(m_armada, name.c_str(), attachment->get()));
        m_attachments[name] = attachment;
        m_active = name;
    }
    void activate(const std::string &name) {
        check(krt_armada_activate(m_armada, name.c_str()));
        m_active = name;
    }
    void set_depth(int depth) { check(krt_armada_set_depth(m_armada, depth)); }

    void render(krt_handle system, const Camera &camera, bool accumulate = false) {
        check(krt_armada_render(m_armada, system, &camera.raw, accumulate ? 1 : 0));
    }
};

// kobra::cuda::hdr_to_ldr (include/cuda/color.cuh:104-118): same arguments,
// plus the context whose stream the launch is queued on.
enum { eTonemappingACES = KRT_TONEMAP_ACES, eTonemappingSRGB = KRT_TONEMAP_SRGB };
inline void hdr_to_ldr(Context &ctx, const void *pixels, uint32_t *target, int width, int height,
                       int tonemapping = eTonemappingACES) {
    check(krt_hdr_to_ldr(ctx.get(), pixels, target, (uint32_t)width, (uint32_t)height, tonemapping));
}

}  // namespace krt_cpp

#endif

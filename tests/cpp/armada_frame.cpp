// Renders through the C++ drop-in layer (include/krt_armada.hpp) on the GPU:
//   argv: scene file, output prefix, width, height, frames,
//         camera px py pz rx ry rz fov aspect
// 1. a PathTracer from load_attachment(), depth 5, `frames` accumulated
//    frames via ArmadaRTX::render(scene, Camera, Transform, accumulate);
//    writes <prefix>_pt.f32 (the colour AOV) and <prefix>_cam.f32 (the
//    krt_camera the layer computed: center, U, V, W, pv).
// 2. a user attachment (a subclass of AttachmentRTX written here) that puts
//    each pixel's closest-hit distance into the colour buffer; writes
//    <prefix>_depth.f32 and the load / unload / render calls it saw.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "krt_armada.hpp"

using namespace krt_cpp;

template <class T>
static std::vector<T> read_array(std::ifstream &in, size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char *>(v.data()), sizeof(T) * n);
    if (!in) throw std::runtime_error("short scene file");
    return v;
}

template <class T>
static void write_file(const std::string &path, const T *p, size_t n) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char *>(p), sizeof(T) * n);
}

// A user attachment: closest-hit distance of the pixel-centre camera ray.
class DepthAttachment : public AttachmentRTX {
public:
    int loads = 0, unloads = 0, renders = 0, attached = 0;
    DepthAttachment() : AttachmentRTX(1) {}
    void attach(const ArmadaRTX &) override { attached++; }
    void load() override { loads++; }
    void unload() override { unloads++; }
    void render(const ArmadaRTX *armada, const ArmadaLaunchInfo &info, const std::optional<krt_handle> &handle,
                std::vector<krt_handle> *, const Extent2D &extent) override {
        renders++;
        krt_camera cam{};
        const Vec3 vs[4] = {info.camera.center, info.camera.ax_u, info.camera.ax_v, info.camera.ax_w};
        float *dst[4] = {cam.center, cam.U, cam.V, cam.W};
        for (int i = 0; i < 4; i++) dst[i][0] = vs[i].x, dst[i][1] = vs[i].y, dst[i][2] = vs[i].z;
        const size_t n = (size_t)extent.width * extent.height;
        std::vector<float> rays(8 * n), hits(4 * n), color(4 * n);
        check(krt_camera_rays(info.context, &cam, extent.width, extent.height, 0, info.samples, 0, rays.data()));
        check(krt_trace_closest(info.context, *handle, rays.data(), (uint32_t)n, hits.data()));
        for (size_t i = 0; i < n; i++) {
            color[4 * i] = color[4 * i + 1] = color[4 * i + 2] = hits[4 * i];
            color[4 * i + 3] = 1.0f;
        }
        check(krt_armada_upload(armada->get(), 0, color.data()));
    }
};

int main(int argc, char **argv) {
    if (argc != 14) {
        std::fprintf(stderr, "usage: scene prefix width height frames px py pz rx ry rz fov aspect\n");
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        uint32_t counts[3];
        in.read(reinterpret_cast<char *>(counts), sizeof(counts));
        const uint32_t nv = counts[0], nt = counts[1], nm = counts[2];
        auto pos = read_array<float>(in, 3 * (size_t)nv);
        auto nrm = read_array<float>(in, 3 * (size_t)nv);
        auto idx = read_array<uint32_t>(in, 3 * (size_t)nt);
        auto tmat = read_array<uint32_t>(in, nt);
        auto mats = read_array<float>(in, 12 * (size_t)nm);
        const std::string prefix = argv[2];
        const Extent2D extent{(uint32_t)std::atoi(argv[3]), (uint32_t)std::atoi(argv[4])};
        const int frames = std::atoi(argv[5]);
        Transform tr;
        tr.position = {(float)std::atof(argv[6]), (float)std::atof(argv[7]), (float)std::atof(argv[8])};
        tr.rotation = {(float)std::atof(argv[9]), (float)std::atof(argv[10]), (float)std::atof(argv[11])};
        Camera camera((float)std::atof(argv[12]), (float)std::atof(argv[13]));

        Context ctx(0);
        Accelerator system(ctx);
        krt_scene_desc d{pos.data(), nrm.data(), nv, idx.data(), tmat.data(), nt, mats.data(), nm};
        const krt_handle scene = system.update(d);
        ArmadaRTX armada(system, extent);
        AttachmentEntry pt = load_attachment(ctx, "path_tracer");
        armada.attach(pt.name, pt.ptr);
        armada.set_depth(5);
        armada.set_option("seed", 0);
        if (std::get<int>(armada.get_option("seed")) != 0) throw std::runtime_error("seed option lost");
        for (int f = 0; f < frames; f++) armada.render(scene, camera, tr, true);
        const auto img = armada.download(0);
        write_file(prefix + "_pt.f32", img.data(), img.size());
        const krt_camera k = make_camera(camera, tr);
        std::vector<float> cam(k.center, k.center + 3);
        cam.insert(cam.end(), k.U, k.U + 3);
        cam.insert(cam.end(), k.V, k.V + 3);
        cam.insert(cam.end(), k.W, k.W + 3);
        cam.insert(cam.end(), k.pv, k.pv + 16);
        write_file(prefix + "_cam.f32", cam.data(), cam.size());

        auto depth = std::make_shared<DepthAttachment>();
        armada.attach("depth", depth);
        armada.render(scene, camera, tr, false);
        const auto dimg = armada.download(0);
        write_file(prefix + "_depth.f32", dimg.data(), dimg.size());
        armada.activate(pt.name);
        armada.render(scene, camera, tr, true);
        armada.activate("depth");
        armada.render(scene, camera, tr, false);
        std::printf("attached %d loads %d unloads %d renders %d\n", depth->attached, depth->loads, depth->unloads,
                    depth->renders);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

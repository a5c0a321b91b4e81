// Prints the C++ layer's camera frames (uvw_frame, Transform axes,
// projection * view) for transforms given on stdin: one line per transform
// "px py pz rx ry rz fov aspect" -> "center U V W pv[16] fwd up".
#include <cstdio>

#include "krt_armada.hpp"

int main() {
    float v[8];
    while (std::scanf("%f %f %f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]) == 8) {
        krt_cpp::Transform t;
        t.position = {v[0], v[1], v[2]};
        t.rotation = {v[3], v[4], v[5]};
        krt_cpp::Camera cam(v[6], v[7]);
        const krt_camera k = krt_cpp::make_camera(cam, t);
        const krt_cpp::Vec3 f = t.forward(), u = t.up();
        for (float x : k.center) std::printf("%.9g ", x);
        for (float x : k.U) std::printf("%.9g ", x);
        for (float x : k.V) std::printf("%.9g ", x);
        for (float x : k.W) std::printf("%.9g ", x);
        for (float x : k.pv) std::printf("%.9g ", x);
        std::printf("%.9g %.9g %.9g %.9g %.9g %.9g\n", f.x, f.y, f.z, u.x, u.y, u.z);
    }
    return 0;
}

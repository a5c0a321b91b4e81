"""Camera frame for a render call (ArmadaLaunchInfo::Camera, armada.cuh:46-55).

`look_at` follows uvw_frame (include/camera.hpp:45-63): W = lookat - eye,
U = normalize(W x up) * tan(fov/2) * |W| * aspect, V = normalize(U x W) *
tan(fov/2) * |W|.  `pv` is projection * view (glm::perspective(fov, aspect,
0.1, 1000) * glm::lookAt, camera.hpp:20-32) in column-major order, used only
for temporal reprojection.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class CameraFrame:
    center: np.ndarray
    U: np.ndarray
    V: np.ndarray
    W: np.ndarray
    pv: np.ndarray  # (4, 4) column-major flattening via .T

    def to_ctypes(self) -> L.Camera:
        c = L.Camera()
        for name in ("center", "U", "V", "W"):
            v = np.asarray(getattr(self, name), np.float32)
            getattr(c, name)[:] = [float(x) for x in v]
        c.pv[:] = [float(x) for x in np.asarray(self.pv, np.float32).reshape(16)]
        return c

    def as_arrays(self):
        return (np.asarray(self.center, np.float32), np.asarray(self.U, np.float32),
                np.asarray(self.V, np.float32), np.asarray(self.W, np.float32),
                np.asarray(self.pv, np.float32).reshape(16))


def _normalize(v):
    return v / np.linalg.norm(v)


def look_at(eye, lookat, up=(0.0, 1.0, 0.0), fov: float = 45.0, aspect: float = 1.0) -> CameraFrame:
    eye = np.asarray(eye, np.float64)
    lookat = np.asarray(lookat, np.float64)
    up = np.asarray(up, np.float64)
    w = lookat - eye
    wlen = np.linalg.norm(w)
    u = _normalize(np.cross(w, up))
    v = _normalize(np.cross(u, w))
    vlen = wlen * np.tan(np.radians(fov) / 2.0)
    v = v * vlen
    u = u * (vlen * aspect)
    # glm::lookAt (right-handed) and glm::perspective (clip z in [-1, 1])
    f = _normalize(w)
    s = _normalize(np.cross(f, up))
    uu = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, uu, -f
    view[0, 3], view[1, 3], view[2, 3] = -s @ eye, -uu @ eye, f @ eye
    t = np.tan(np.radians(fov) / 2.0)
    near, far = 0.1, 1000.0
    proj = np.zeros((4, 4))
    proj[0, 0] = 1.0 / (aspect * t)
    proj[1, 1] = 1.0 / t
    proj[2, 2] = -(far + near) / (far - near)
    proj[2, 3] = -2.0 * far * near / (far - near)
    proj[3, 2] = -1.0
    pv = (proj @ view).astype(np.float32)
    return CameraFrame(eye.astype(np.float32), u.astype(np.float32), v.astype(np.float32),
                       w.astype(np.float32), pv.T.copy().reshape(16))

"""The C++ drop-in layer (include/krt_armada.hpp) on the host: it compiles
against include/krt.h and links to libkrt.so, and its camera frame --
Transform::forward/up (include/transform.hpp:96-111, degrees passed to
glm::quat unconverted), uvw_frame (include/camera.hpp:45-63) and
perspective * lookAt (camera.hpp:20-32) -- equals the Python layer's
(artemis_amd.project.transform_axes + look_at, whose axis conventions are
pinned against scipy in tests/test_project_loader.py) to float rounding."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from artemis_amd import look_at
from artemis_amd.project import transform_axes

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "artemis_amd" / "libkrt.so"


def compile_cpp(tmp_path, name):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    if not LIB.exists():
        pytest.fail("libkrt.so is not built")
    exe = tmp_path / name
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                        str(ROOT / "tests" / "cpp" / f"{name}.cpp"), str(LIB), f"-Wl,-rpath,{LIB.parent}", "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_camera_frame_matches_python_layer(tmp_path):
    exe = compile_cpp(tmp_path, "camera_math")
    rng = np.random.default_rng(4)
    cases = [(0.0, 1.0, 3.5, 0.0, 0.0, 0.0, 45.0, 1.0)]
    for _ in range(40):
        p = rng.uniform(-10, 10, 3)
        r = rng.uniform(-3, 3, 3)
        cases.append((*p, *r, float(rng.uniform(20, 90)), float(rng.uniform(0.5, 2.0))))
    cases = [tuple(float(np.float32(v)) for v in c) for c in cases]  # exact round trip through %.9g
    stdin = "\n".join(" ".join(f"{v:.9g}" for v in c) for c in cases) + "\n"
    out = subprocess.run([str(exe)], input=stdin, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    rows = [np.array(line.split(), np.float64) for line in out.stdout.strip().splitlines()]
    assert len(rows) == len(cases)
    for c, row in zip(cases, rows):
        pos = np.array(c[:3], np.float32)
        fwd, up = transform_axes(c[3:6])
        np.testing.assert_allclose(row[28:31], fwd, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(row[31:34], up, rtol=1e-5, atol=1e-6)
        frame = look_at(pos, pos + fwd, up=up, fov=c[6], aspect=c[7])
        center, U, V, W, pv = frame.as_arrays()
        np.testing.assert_array_equal(row[0:3].astype(np.float32), center)
        for got, want in ((row[3:6], U), (row[6:9], V), (row[9:12], W)):
            np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(row[12:28], pv, rtol=2e-4, atol=2e-4)


def test_cpp_drop_in_layer_links(tmp_path):
    exe = compile_cpp(tmp_path, "armada_frame")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and "usage" in out.stderr  # argument check before any HIP call

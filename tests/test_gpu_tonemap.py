"""hdr_to_ldr on the GPU (krt_hdr_to_ldr, krt_armada_read_ldr) vs the oracle
restatement of include/cuda/color.cuh:82-121.

ACES is plain IEEE arithmetic in the reference's order, so the codes must be
bit-exact.  sRGB calls powf, whose last bit may differ between the device
library and glibc; codes may then differ by one step, and only in a small
fraction of pixels (the reference's own CUDA powf is not reproduced either)."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, PathTracer, look_at, scenes

pytestmark = pytest.mark.gpu


def hdr_pixels(H, W, seed):
    rng = np.random.default_rng(seed)
    px = (rng.random((H, W, 4)) ** 3 * 12 - 0.5).astype(np.float32)
    flat = px.reshape(-1, 4)
    specials = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, 1e-8, 0.0031308, 1e30], np.float32)
    n = min(len(specials), len(flat))
    flat[:n, 0] = specials[:n]
    return px


@pytest.mark.parametrize("W,H", [(1, 1), (257, 3), (640, 360), (1920, 1080)])
@pytest.mark.parametrize("mode", [0, 1])
def test_device_matches_oracle(krt_ctx, W, H, mode):
    px = hdr_pixels(H, W, W * 7 + H + mode)
    src = krt_ctx.device_array(px)
    dst = krt_ctx.device_array(shape=(H, W), dtype=np.uint32)
    krt_ctx.hdr_to_ldr(src.ptr, dst.ptr, W, H, mode)
    got = dst.download().reshape(H, W)
    want = O.hdr_to_ldr(px, mode)
    if mode == 0:
        assert np.array_equal(got, want)
    else:
        diff = np.zeros(got.shape, np.int64)
        for ch in range(4):
            d = ((got >> (8 * ch)) & 255).astype(np.int64) - ((want >> (8 * ch)) & 255).astype(np.int64)
            diff = np.maximum(diff, np.abs(d))
        assert diff.max() <= 1
        assert np.mean(diff == 0) >= 0.999


def test_empty_image_is_a_no_op(krt_ctx):
    krt_ctx.hdr_to_ldr(0, 0, 0, 0, 0)
    krt_ctx.sync()


def test_bad_mode_is_rejected(krt_ctx):
    src = krt_ctx.device_array(shape=(4, 4, 4))
    dst = krt_ctx.device_array(shape=(4, 4), dtype=np.uint32)
    with pytest.raises(RuntimeError):
        krt_ctx.hdr_to_ldr(src.ptr, dst.ptr, 4, 4, 2)


def test_armada_ldr_buffer(krt_ctx):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    W, H = 96, 64
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("pt", PathTracer(krt_ctx))
    cam = look_at(**s.camera, aspect=W / H)
    for _ in range(2):
        arm.render(sc, cam)
    hdr = arm.color_buffer()
    assert np.array_equal(arm.ldr_buffer(0), O.hdr_to_ldr(hdr, 0))
    img = arm.ldr_buffer(0)
    assert (img >> 24 == 255).all() and np.any(img & 255)
    sc.free()


def test_armada_png_file(krt_ctx, tmp_path):
    """krt_armada_write_png: the colour buffer through hdr_to_ldr into a PNG
    file whose pixels are the LDR buffer's bytes."""
    from .test_png_output import read_png

    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    arm = ArmadaRTX(krt_ctx, 40, 24)
    arm.attach("pt", PathTracer(krt_ctx))
    arm.render(sc, look_at(**s.camera, aspect=40 / 24))
    path = tmp_path / "frame.png"
    arm.write_png(path, 0)
    ldr = arm.ldr_buffer(0)
    np.testing.assert_array_equal(read_png(path), ldr.view(np.uint8).reshape(24, 40, 4))
    sc.free()

"""The kernels' reprojection index rounding (round_half_up in
artemis_amd/csrc/krt_device.h: floor(x) + (x - floor(x) >= 0.5), all in
FP32) equals the reference's (int)((double)x + 0.5)
(source/optix/amadeus_restir.cu:202-203) for every pixel coordinate: checked
on every float within 64 ulps of each half-integer and integer below 4,096
and on 2^22 random floats, with numpy's float32 arithmetic standing in for
the kernel's."""
import numpy as np


def fp32_form(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.float32)
    i = x.astype(np.int32)  # truncation == floor for x >= 0
    frac = (x - i.astype(np.float32)).astype(np.float32)
    return i + (frac >= np.float32(0.5)).astype(np.int32)


def fp64_form(x: np.ndarray) -> np.ndarray:
    return (x.astype(np.float64) + 0.5).astype(np.int32)


def test_round_half_up_matches_double_form():
    centres = np.concatenate([np.arange(0, 4096, dtype=np.float32) + np.float32(0.5),
                              np.arange(0, 4097, dtype=np.float32)])
    vals = [centres]
    up, down = centres.copy(), centres.copy()
    for _ in range(64):
        up = np.nextafter(up, np.float32(np.inf))
        down = np.nextafter(down, np.float32(-np.inf))
        vals += [up, down]
    x = np.concatenate(vals)
    x = x[x >= 0]
    rng = np.random.default_rng(3)
    x = np.concatenate([x, (rng.random(1 << 22) * 4096).astype(np.float32), np.float32([0.0, 1920.0, 1080.0])])
    np.testing.assert_array_equal(fp32_form(x), fp64_form(x))
    # a plain float sum would not: 0.5 - 2^-25 + 0.5 rounds to 1.0 in float
    y = np.float32(0.5) - np.float32(2.0 ** -25)
    assert fp64_form(np.array([y]))[0] == 0 and int(np.float32(y + np.float32(0.5))) == 1

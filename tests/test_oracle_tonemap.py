"""Oracle hdr_to_ldr (include/cuda/color.cuh:14-121) against closed forms.

The expected codes come from the published curves evaluated in exact
rational (ACES, Narkowicz 2015) or float64 (sRGB, IEC 61966-2-1) arithmetic,
then floor(256 v) capped at 255; inputs are chosen away from code
boundaries, so float32 rounding cannot move them.  The packing (r in the low
byte, alpha 255, include/cuda/cast.cuh:16-20), the vertical flip
(color.cuh:92-93) and the clamp quirk (fmaxf(a, fminf(f, b)),
include/cuda/math.cuh:101-104: NaN -> 1) are checked directly.
"""
from fractions import Fraction

import numpy as np

import oracle as O


def aces_exact(c: float) -> Fraction:
    x = Fraction(c)
    v = x * (Fraction(251, 100) * x + Fraction(3, 100)) / (x * (Fraction(243, 100) * x + Fraction(59, 100)) + Fraction(14, 100))
    return min(max(v, Fraction(0)), Fraction(1))


def code(v) -> int:
    return min(int(v * 256), 255)


def srgb_f64(c: float) -> float:
    s = 12.92 * c if c < 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055
    return min(max(s, 0.0), 1.0)


def one(c, mode):
    px = np.zeros((1, 1, 4), np.float32)
    px[0, 0, :3] = c
    v = int(O.hdr_to_ldr(px, mode)[0, 0])
    return v & 255, (v >> 8) & 255, (v >> 16) & 255, v >> 24


def test_aces_closed_form():
    for c in [0.0, 0.01, 0.1, 0.18, 0.5, 1.0, 2.0, 4.0, 16.0, 1e6]:
        v = aces_exact(np.float32(c).item())
        frac = (v * 256) - int(v * 256)
        if 0 < v < 1:
            assert 1e-4 < frac < 1 - 1e-4, c  # keep the case away from a code boundary
        r, g, b, a = one(c, 0)
        assert (r, g, b, a) == (code(v),) * 3 + (255,), c
    assert one(1.0, 0)[0] == 205  # 2.54 / 3.16 = 0.8038 -> 205


def test_srgb_closed_form():
    for c in [0.0, 0.001, 0.0031, 0.01, 0.05, 0.2, 0.5, 0.9, 1.0, 3.0]:
        v = srgb_f64(float(np.float32(c)))
        if v < 1 - 1e-9:
            frac = v * 256 - int(v * 256)
            assert 1e-3 < frac < 1 - 1e-3 or v == 0.0, c
        assert one(c, 1)[:3] == (code(v),) * 3, c
    assert one(0.5, 1)[0] == 188  # 1.055 * 0.5^(1/2.4) - 0.055 = 0.7354
    assert one(0.001, 1)[0] == 3  # linear segment: 12.92e-3 * 256 = 3.3


def test_special_values_follow_the_reference_clamp():
    nan, inf = float("nan"), float("inf")
    assert one(nan, 0)[0] == 255 and one(nan, 1)[0] == 255  # fminf(NaN, 1) = 1
    # the ACES rational maps -1 to 2.48 / 1.98 > 1 (no clamp before the curve)
    assert one(-1.0, 0)[0] == code(aces_exact(-1.0)) == 255
    assert one(-1.0, 1)[0] == 0
    assert one(inf, 0)[0] == 255 and one(inf, 1)[0] == 255
    assert one(-0.0, 0)[0] == 0


def test_packing_and_flip():
    H, W = 3, 5
    px = np.zeros((H, W, 4), np.float32)
    px[0, 0] = [1e6, 0.0, 0.0, 7.0]  # red in the first input row
    px[H - 1, W - 1] = [0.0, 0.0, 1e6, 0.0]
    out = O.hdr_to_ldr(px, 0)
    assert out.shape == (H, W)
    assert out[H - 1, 0] == 0xFF0000FF  # r = 255 in the low byte, alpha 255, last output row
    assert out[0, W - 1] == 0xFFFF0000
    assert np.all(out[1] == 0xFF000000)


def test_random_pixels_within_one_code_of_float64():
    rng = np.random.default_rng(3)
    px = (rng.random((64, 48, 4)) ** 3 * 8).astype(np.float32)
    for mode, f in ((0, lambda c: float(aces_exact(c))), (1, srgb_f64)):
        out = O.hdr_to_ldr(px, mode)
        flipped = out[::-1]
        for ch in range(3):
            got = (flipped >> (8 * ch)) & 255
            want = np.vectorize(lambda c: code(f(float(c))))(px[..., ch])
            assert np.abs(got.astype(int) - want).max() <= 1
            assert np.mean(got == want) > 0.99

"""krt_read_png (the image input behind material textures) against an
independent PNG encoder written here: every colour type and bit depth, all
five scanline filters chosen at random per row, Adam7 interlacing at odd
sizes, palettes with tRNS alpha and grey / RGB transparency keys; the
expected RGBA8 follows the PNG specification (16-bit samples keep their high
byte, low-depth grey scales to 0..255).  Also: krt_write_png's files read
back bit for bit, and broken files are refused.  Host code only."""
import struct
import zlib

import numpy as np
import pytest

from artemis_amd import read_png, write_png
from artemis_amd._lib import KRT_ERR_INVALID, KrtError

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def pack_row(samples, depth):
    """Samples of one row (already interleaved channels) -> bytes."""
    if depth == 16:
        return np.asarray(samples, ">u2").tobytes()
    if depth == 8:
        return np.asarray(samples, np.uint8).tobytes()
    per = 8 // depth
    out = bytearray((len(samples) + per - 1) // per)
    for i, v in enumerate(samples):
        out[i // per] |= int(v) << (8 - depth - (i % per) * depth)
    return bytes(out)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def filter_rows(rows, bpp, rng):
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for row in rows:
        ft = int(rng.integers(0, 5))
        f = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][ft]
            f[i] = (row[i] - pred) & 255
        out += bytes([ft]) + f
        prev = row
    return bytes(out)


def encode(samples, ctype, depth, interlace, rng, plte=None, trns=None):
    """samples: (H, W, channels) integers of the file's depth."""
    H, W, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        raw += filter_rows(rows, bpp, rng)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, interlace))
    if plte is not None:
        png += chunk(b"PLTE", np.asarray(plte, np.uint8).tobytes())
    if trns is not None:
        png += chunk(b"tRNS", trns)
    png += chunk(b"tEXt", b"Comment\x00ancillary chunk")  # skipped by the reader
    half = len(zlib.compress(raw)) // 2
    z = zlib.compress(raw)
    png += chunk(b"IDAT", z[:half]) + chunk(b"IDAT", z[half:])  # split IDAT
    return png + chunk(b"IEND", b"")


CASES = [(0, d) for d in (1, 2, 4, 8, 16)] + [(2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (4, 16),
                                               (6, 8), (6, 16)]


@pytest.mark.parametrize("interlace", [0, 1])
@pytest.mark.parametrize("ctype,depth", CASES)
def test_decoder_matches_the_specification(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth * 2 + interlace)
    H, W = 11, 13
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    samples = rng.integers(0, top + 1, (H, W, ch), dtype=np.int64)
    plte = trns = None
    want = np.zeros((H, W, 4), np.uint8)
    if ctype == 3:
        n = top + 1
        plte = rng.integers(0, 256, (n, 3))
        alpha = rng.integers(0, 256, n // 2 + 1)  # entries past tRNS stay opaque
        trns = alpha.astype(np.uint8).tobytes()
        idx = samples[..., 0]
        want[..., :3] = plte[idx]
        full = np.full(n, 255)
        full[:len(alpha)] = alpha
        want[..., 3] = full[idx]
    else:
        to8 = (samples >> 8) if depth == 16 else (samples * 255 // top if depth < 8 else samples)
        if ctype in (0, 4):
            want[..., 0] = want[..., 1] = want[..., 2] = to8[..., 0]
            want[..., 3] = to8[..., 1] if ctype == 4 else 255
        else:
            want[..., :3] = to8[..., :3]
            want[..., 3] = to8[..., 3] if ctype == 6 else 255
        if ctype in (0, 2):  # a transparency key: one sample value becomes alpha 0
            key = samples[3, 5, :ch]
            trns = b"".join(struct.pack(">H", int(v)) for v in key)
            want[np.all(samples[..., :ch] == key, axis=-1), 3] = 0
    path = tmp_path / f"t{ctype}_{depth}_{interlace}.png"
    path.write_bytes(encode(samples, ctype, depth, interlace, rng, plte, trns))
    got = read_png(path)
    assert got.shape == (H, W, 4)
    assert np.array_equal(got, want)


def test_written_files_read_back(tmp_path):
    rng = np.random.default_rng(9)
    px = rng.integers(0, 2 ** 32, (37, 53), dtype=np.uint64).astype(np.uint32)
    write_png(tmp_path / "a.png", px)
    assert np.array_equal(read_png(tmp_path / "a.png").view(np.uint32)[..., 0], px)


def test_broken_files_are_refused(tmp_path):
    rng = np.random.default_rng(2)
    good = encode(rng.integers(0, 256, (4, 4, 3)), 2, 8, 0, rng)
    bad_crc = bytearray(good)
    bad_crc[20] ^= 1  # inside IHDR
    for name, data in (("truncated", good[:40]), ("crc", bytes(bad_crc)), ("not_png", b"GIF89a" + good[6:])):
        p = tmp_path / f"{name}.png"
        p.write_bytes(data)
        with pytest.raises(KrtError) as e:
            read_png(p)
        assert e.value.code == KRT_ERR_INVALID, name
    with pytest.raises(KrtError):
        read_png(tmp_path / "missing.png")

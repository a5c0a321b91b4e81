"""PNG output (krt_write_png, the reference's stbi_write_png still capture of
the hdr_to_ldr frame, bench.cu:226-233), decoded here with the PNG
specification's rules (chunk layout, CRC-32, zlib stream, filter-0 rows)
from Python's zlib: the file holds exactly the packed RGBA8 pixels."""
import struct
import zlib

import numpy as np
import pytest

from artemis_amd import _lib as L, write_png


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (64, 48)])
def test_png_holds_the_pixels(tmp_path, w, h):
    rng = np.random.default_rng(w * h)
    px = rng.integers(0, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)
    path = tmp_path / "x.png"
    write_png(path, px)
    got = read_png(path)
    want = px.view(np.uint8).reshape(h, w, 4)  # little-endian r | g << 8 | b << 16 | a << 24
    np.testing.assert_array_equal(got, want)


def test_bad_arguments_are_refused(tmp_path):
    with pytest.raises(L.KrtError):
        write_png(tmp_path / "missing_dir" / "x.png", np.zeros((2, 2), np.uint32))
    lib = L.lib()
    assert lib.krt_write_png(None, None, 1, 1) == L.KRT_ERR_INVALID

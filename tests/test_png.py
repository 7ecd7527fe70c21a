"""The frame file writer (ifl_write_png, standing in for the reference's
lodepng_encode32_file, 7:1122): host logic, no device. The file is decoded
here from the PNG and zlib specifications (RFC 2083 / 1950 / 1951) with
Python's zlib and checked chunk by chunk."""
import struct
import zlib

import numpy as np


def decode(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) == crc, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # also checks the Adler-32
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip(tmp_path):
    import ifl_fluids as F
    rng = np.random.default_rng(3)
    for w, h in ((1, 1), (7, 5), (300, 220)):  # 300x220 spans several stored blocks
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        p = tmp_path / f"f{w}x{h}.png"
        F.write_png(str(p), img)
        assert np.array_equal(decode(str(p)), img)


def test_png_rejects_bad_arguments(tmp_path):
    import ctypes
    import ifl_fluids as F
    buf = (ctypes.c_uint8 * 16)()
    assert F.lib.ifl_write_png(str(tmp_path / "x.png").encode(), buf, 0, 1) != 0
    assert F.lib.ifl_write_png(b"/nonexistent-dir/x.png", buf, 2, 2) != 0

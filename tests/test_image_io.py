"""Render output files (CPU): the PFM float dump and the PNG of the
reference's displayed bytes (pic2RGBW, CudaPathTracing/Render.cu:462-481;
NORMAL / DEPTH display, Render.cu:239-258).  Each file is parsed back here
with Python's own zlib and CRC-32, independently of the writer."""
import struct
import zlib

import numpy as np
import pytest

import cudapathtracing_amd as cpt


def image(h=37, w=29, seed=0):
    rng = np.random.default_rng(seed)
    rgb = rng.uniform(-0.2, 1.6, (h, w, 3)).astype(np.float32)
    rgb[0, 0] = (np.nan, np.inf, -np.inf)
    rgb[1, 1] = (0.0, 1.0, 0.25)
    return rgb


def test_pfm_round_trip(tmp_path):
    rgb = image()
    p = tmp_path / "a.pfm"
    cpt.write_pfm(str(p), rgb)
    data = p.read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"29 37" and float(head[2]) == -1.0
    back = np.frombuffer(head[3], "<f4").reshape(37, 29, 3)  # rows bottom to top, as stored
    assert np.array_equal(back, rgb, equal_nan=True)


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xffffffff, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # checks the Adler-32 as well
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("gamma", [True, False])
def test_png_matches_display_bytes(tmp_path, gamma):
    rgb = image()
    p = tmp_path / "a.png"
    cpt.write_png(str(p), rgb, gamma=gamma)
    px = read_png(str(p))
    # the display bytes: uchar(254.99 * clamp01(sqrt(x))) (gamma) or of x;
    # NaN shows 0; the PNG's first row is the image's top row (row h - 1)
    x = rgb.astype(np.float64)
    with np.errstate(invalid="ignore"):
        if gamma:
            x = np.sqrt(x)
    want = np.floor(254.99 * np.nan_to_num(np.clip(x, 0.0, 1.0), nan=0.0)).astype(np.uint8)
    assert np.array_equal(px[..., :3], want[::-1])
    assert np.all(px[..., 3] == 255)
    if gamma:
        assert np.array_equal(px[..., :3], cpt.tonemap_rgba8(rgb)[::-1, :, :3])


def test_png_large_image_uses_several_stored_blocks(tmp_path):
    rgb = image(300, 300, seed=1)  # 360,300 raw bytes > 65,535 per stored block
    p = tmp_path / "b.png"
    cpt.write_png(str(p), rgb)
    assert read_png(str(p)).shape == (300, 300, 4)


def test_unwritable_path_is_an_io_error(tmp_path):
    with pytest.raises(cpt.CptError) as e:
        cpt.write_png(str(tmp_path / "missing" / "x.png"), image())
    assert e.value.code == -5

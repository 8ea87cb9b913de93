"""PNG input and output (rts_png_*), the replacement of the reference
driver's OpenCV imread / imwrite (seq/Main.cpp:72-73, 87-88, 275-282).  CPU
only: files written by PIL (an independent encoder) in every colour type and
bit depth it produces, an Adam7-interlaced file assembled here from the
specification, round trips, PIL decoding our files, damaged files, and the
display scaling of seq/Main.cpp:335-360 on closed-form cases."""
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from rtstereo_amd import gen
from rtstereo_amd import io as rio
from rtstereo_amd._lib import RtsError


def _bgr(rgb):
    return np.ascontiguousarray(rgb[..., ::-1])


@pytest.mark.parametrize("W,H", [(1, 1), (7, 5), (64, 48), (129, 33)])
def test_round_trip(tmp_path, W, H):
    rng = np.random.default_rng(W * 100 + H)
    g = rng.integers(0, 256, size=(H, W), dtype=np.uint8)
    c = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8)
    smooth = gen.make_pair(max(W, 16), max(H, 16), 8, 2, 1.0, 3).left[:H, :W]
    for img in (g, c, np.ascontiguousarray(smooth)):
        p = str(tmp_path / "x.png")
        rio.write_png(p, img)
        back = rio.read_png(p, 1 if img.ndim == 2 else 3)
        assert np.array_equal(back, img)
        # an independent decoder reads the same pixels (PIL gives R,G,B)
        pil = np.asarray(Image.open(p))
        assert np.array_equal(pil, img if img.ndim == 2 else _bgr(img))


def test_pil_written_files(tmp_path):
    rng = np.random.default_rng(5)
    H, W = 37, 53
    rgb = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8)
    rgba = rng.integers(0, 256, size=(H, W, 4), dtype=np.uint8)
    gray = rng.integers(0, 256, size=(H, W), dtype=np.uint8)
    cases = {
        "rgb": (Image.fromarray(rgb, "RGB"), {}, rgb),
        "rgb_opt": (Image.fromarray(rgb, "RGB"), {"optimize": True}, rgb),
        "rgb_nocomp": (Image.fromarray(rgb, "RGB"), {"compress_level": 0}, rgb),
        "rgba": (Image.fromarray(rgba, "RGBA"), {}, rgba[..., :3]),
        "gray": (Image.fromarray(gray, "L"), {}, np.stack([gray] * 3, -1)),
        "gray_alpha": (Image.fromarray(np.stack([gray, gray[::-1]], -1), "LA"), {}, np.stack([gray] * 3, -1)),
    }
    pal = Image.fromarray(rgb, "RGB").quantize(colors=200)
    pal_rgb = np.asarray(pal.convert("RGB"))
    cases["palette8"] = (pal, {}, pal_rgb)
    pal16 = Image.fromarray(rgb, "RGB").quantize(colors=16)
    cases["palette4"] = (pal16, {"bits": 4}, np.asarray(pal16.convert("RGB")))
    bw = Image.fromarray((gray > 127).astype(np.uint8) * 255, "L").convert("1")
    cases["bilevel"] = (bw, {}, np.stack([np.asarray(bw.convert("L"))] * 3, -1))
    g16 = (gray.astype(np.uint16) * 257 + rng.integers(0, 256, size=(H, W)).astype(np.uint16))
    cases["gray16"] = (Image.fromarray(g16), {}, np.stack([(g16 >> 8).astype(np.uint8)] * 3, -1))
    for name, (im, opts, want_rgb) in cases.items():
        p = str(tmp_path / f"{name}.png")
        im.save(p, **opts)
        got = rio.read_png(p, 3)
        assert np.array_equal(got, _bgr(want_rgb)), name
        g = rio.read_png(p, 1)
        if im.mode in ("L", "LA", "1", "I;16", "I"):
            assert np.array_equal(g, want_rgb[..., 0]), name
        else:
            assert np.array_equal(g, gen.to_gray(_bgr(want_rgb))), name
        with open(p, "rb") as f:
            assert np.array_equal(rio.decode_png(f.read(), 3), got), name


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def test_adam7_interlaced():
    """An interlaced RGB file assembled from the specification's seven
    passes, each scanline with a different filter."""
    rng = np.random.default_rng(9)
    H, W = 13, 19
    rgb = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    raw = b""
    k = 0
    for sx, sy, dx, dy in passes:
        sub = rgb[sy::dy, sx::dx]
        if sub.size == 0:
            continue
        prev = np.zeros(sub.shape[1] * 3, np.int32)
        for row in sub:
            cur = row.reshape(-1).astype(np.int32)
            ft = k % 3   # None, Sub, Up
            k += 1
            if ft == 0:
                f = cur
            elif ft == 1:
                f = cur - np.concatenate([np.zeros(3, np.int32), cur[:-3]])
            else:
                f = cur - prev
            raw += bytes([ft]) + (f & 255).astype(np.uint8).tobytes()
            prev = cur
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 1))
    z = zlib.compress(raw)
    png += _chunk(b"IDAT", z[:40]) + _chunk(b"IDAT", z[40:]) + _chunk(b"IEND", b"")
    assert np.array_equal(rio.decode_png(png, 3), _bgr(rgb))
    assert np.array_equal(np.asarray(Image.open(__import__("io").BytesIO(png))), rgb)


def test_damaged_files(tmp_path):
    p = str(tmp_path / "ok.png")
    rio.write_png(p, np.arange(48, dtype=np.uint8).reshape(6, 8))
    data = open(p, "rb").read()
    with pytest.raises(RtsError):
        rio.decode_png(b"GIF89a" + data[6:], 1)                 # signature
    bad = bytearray(data)
    bad[40] ^= 0xFF                                                # inside IDAT: CRC mismatch
    with pytest.raises(RtsError):
        rio.decode_png(bytes(bad), 1)
    with pytest.raises(RtsError):
        rio.decode_png(data[:len(data) - 20], 1)                   # truncated
    with pytest.raises(RtsError):
        rio.read_png(str(tmp_path / "missing.png"), 1)


def test_disp_scale_closed_forms():
    d = np.arange(16, dtype=np.uint8).reshape(4, 4)
    assert np.array_equal(rio.disp_scale(d, 16.0), (d.astype(np.int32) * 16).clip(0, 255).astype(np.uint8))
    assert np.array_equal(rio.disp_scale(np.array([[20, 15]], np.uint8), 16.0), np.array([[255, 240]], np.uint8))
    # automatic: the maximum becomes 255; 1 * 127.5 rounds half up to 128
    assert np.array_equal(rio.disp_scale(np.array([[0, 1, 2]], np.uint8)), np.array([[0, 128, 255]], np.uint8))
    assert np.array_equal(rio.disp_scale(np.zeros((2, 3), np.uint8)), np.zeros((2, 3), np.uint8))

"""Film output (gy_image_write / gy_film_save): Film::SaveImage
(visual/Film.cpp:6-9) through ImageIO::WriteImage (image/ImageIO.cpp:120-149).

- The reference writes with stb_image_write (external/stb). oracle/Makefile
  builds that writer from the reference tree into oracle/_ref/img_ref.
- BMP files must equal its output byte for byte. HDR files must equal it
  after the header, whose comment line differs.
- PNG pixels must equal it once decoded: its deflate encoder and filters are
  not reproduced.
- gy_image_read (Radiance .hdr, the loader EnvironmentMap(filename) uses
  through stbi_loadf) must return the floats of the reference's own stb_image
  (external/stb/stb_image.h, built by oracle/Makefile into oracle/_ref/hdr_ref)
  bit for bit: files from both writers, flat small-width files, long runs,
  the non-run-length fallback and header variants.
- These comparisons skip where the reference tree is absent.
- The 8-bit mapping ImageIO::toInt is checked against closed forms.
"""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import gyunity_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IMG_REF = os.path.join(ROOT, "oracle", "_ref", "img_ref")
HDR_REF = os.path.join(ROOT, "oracle", "_ref", "hdr_ref")


def _image(w, h, seed):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0, 1.3, (h, w, 3)).astype(np.float32) ** 3
    img[0, :3] = 0                      # black: RGBE zero
    img[min(1, h - 1), :] = img[min(1, h - 1), 0]  # a constant row: runs in every plane
    img[h - 1, w - 1] = (5e3, 1e-40, 0)  # large and denormal values
    return img


def _ref_write(path, w, h, payload):
    raw = path + ".raw"
    with open(raw, "wb") as f:
        f.write(struct.pack("<ii", w, h))
        f.write(payload.tobytes())
    subprocess.check_call([IMG_REF, raw, path])
    return open(path, "rb").read()


def _need_ref():
    if not os.path.exists(IMG_REF):
        pytest.skip("oracle/_ref/img_ref is built only where the reference tree is present")


def _png_pixels(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", None, None
    while pos < len(data):
        n, t = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert zlib.crc32(t + body) & 0xFFFFFFFF == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        if t == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif t == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = w * 3
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        row = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - 3] if i >= 3 else 0
            b = prev[i]
            c = prev[i - 3] if i >= 3 else 0
            if f == 0:
                pr = 0
            elif f == 1:
                pr = a
            elif f == 2:
                pr = b
            elif f == 3:
                pr = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            cur[i] = (row[i] + pr) & 255
        out[y], prev = cur, cur
    return out.reshape(h, w, 3).astype(np.uint8)


def test_to_int_closed_forms():
    x = np.array([0, 1, -3, 7, 0.5, 1e-30, np.nan, 0.2], np.float32)
    b = g.image_to_bytes(x)
    assert list(b[:4]) == [0, 255, 0, 255]
    assert b[4] == int(0.5 ** (1 / 2.2) * 255 + 0.5)  # 186
    assert b[5] == 0 and b[6] == 0
    # monotone over [0, 1], and every byte value is reached
    v = g.image_to_bytes(np.linspace(0, 1, 200001, dtype=np.float32))
    assert np.all(np.diff(v.astype(np.int32)) >= 0) and len(np.unique(v)) == 256
    # away from the rounding points it equals the formula in double
    xs = np.random.default_rng(1).uniform(0, 1, 20000).astype(np.float32)
    ref = np.power(xs.astype(np.float64), np.float64(np.float32(1) / np.float32(2.2))) * 255 + 0.5
    safe = np.abs(ref - np.round(ref)) > 1e-3
    assert np.array_equal(g.image_to_bytes(xs)[safe], np.floor(ref[safe]).astype(np.uint8))


@pytest.mark.parametrize("w,h", [(37, 11), (64, 3), (5, 4)])
def test_bmp_equals_reference_writer(tmp_path, w, h):
    _need_ref()
    img = _image(w, h, 2)
    ours = str(tmp_path / "ours.bmp")
    g.image_write(ours, img)
    ref = _ref_write(str(tmp_path / "ref.bmp"), w, h, g.image_to_bytes(img))
    assert open(ours, "rb").read() == ref


@pytest.mark.parametrize("w,h", [(300, 7), (9, 5), (5, 4), (200, 1)])
def test_hdr_equals_reference_writer(tmp_path, w, h):
    _need_ref()
    img = _image(w, h, 3)
    img[2 % h, : w // 2] = img[2 % h, 0]  # long runs (> 127)
    ours = str(tmp_path / "ours.hdr")
    g.image_write(ours, img)
    ref = _ref_write(str(tmp_path / "ref.hdr"), w, h, img)
    a, b = open(ours, "rb").read(), ref
    tail = ("-Y %d +X %d\n" % (h, w)).encode()
    assert a[a.index(tail) + len(tail):] == b[b.index(tail) + len(tail):]
    assert b"FORMAT=32-bit_rle_rgbe\n" in a


def test_png_pixels_equal_reference_writer(tmp_path):
    _need_ref()
    img = _image(45, 17, 4)
    ours = str(tmp_path / "ours.png")
    g.image_write(ours, img)
    ref = _ref_write(str(tmp_path / "ref.png"), 45, 17, g.image_to_bytes(img))
    assert np.array_equal(_png_pixels(open(ours, "rb").read()), _png_pixels(ref))
    assert np.array_equal(_png_pixels(open(ours, "rb").read()), g.image_to_bytes(img))


def test_image_write_rejects_unknown_suffix(tmp_path):
    with pytest.raises(g.GyError):
        g.image_write(str(tmp_path / "x.jpg"), np.zeros((2, 2, 3), np.float32))


def _ref_read(path, tmp):
    out = os.path.join(tmp, "read.raw")
    subprocess.check_call([HDR_REF, path, out])
    raw = open(out, "rb").read()
    w, h, c = struct.unpack("<iii", raw[:12])
    return np.frombuffer(raw, np.float32, w * h * c, 12).reshape(h, w, c)


def _need_hdr_ref():
    if not os.path.exists(HDR_REF):
        pytest.skip("oracle/_ref/hdr_ref is built only where the reference tree is present")


def _rgbe(rgb):
    """RGBE bytes of float pixels (stb_image_write's conversion)."""
    out = []
    for r, gch, b in rgb.reshape(-1, 3):
        m = max(r, gch, b)
        if m < 1e-32:
            out.append((0, 0, 0, 0))
            continue
        mant, ex = np.frexp(np.float32(m))
        sc = np.float32(mant * np.float32(256.0) / np.float32(m))
        out.append((int(np.float32(r) * sc), int(np.float32(gch) * sc), int(np.float32(b) * sc), ex + 128))
    return np.array(out, np.uint8)


@pytest.mark.parametrize("w,h", [(37, 11), (5, 4), (300, 3)])
def test_hdr_read_equals_reference_loader(tmp_path, w, h):
    _need_hdr_ref()
    img = _image(w, h, 5)
    img[1 % h, : w // 2] = img[1 % h, 0]  # runs
    ours = str(tmp_path / "ours.hdr")
    g.image_write(ours, img)
    got = g.image_read(ours)
    assert got.shape == (h, w, 3)
    assert np.array_equal(got, _ref_read(ours, str(tmp_path)))
    # 8-bit mantissas: within 2^-8 of the largest channel
    m = img.max(axis=2, keepdims=True)
    assert np.all(np.abs(got - img) <= m * 2.0 ** -7 + 1e-30)
    if os.path.exists(IMG_REF):  # the reference writer's file reads the same
        ref_file = str(tmp_path / "ref.hdr")
        _ref_write(ref_file, w, h, img)
        assert np.array_equal(g.image_read(ref_file), got)


def test_hdr_read_crafted_files(tmp_path):
    _need_hdr_ref()
    rng = np.random.default_rng(9)
    w, h = 200, 3
    pix = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    pix[..., 3] = rng.integers(100, 150, (h, w))
    pix[0, 10:190] = pix[0, 10]  # a run longer than 127
    pix[2, :, 3] = 0             # exponent 0: zero pixels
    pix[0, 0, 0] = 7             # the first pixel does not open with 2 2
    body, first_line = b"", 0
    for j in range(h):
        if j == 1:
            first_line = len(body)
        body += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            row = pix[j, :, k]
            i = 0
            while i < w:
                r = i
                while r < w and r - i < 127 and row[r] == row[i]:
                    r += 1
                if r - i >= 3:
                    body += bytes([128 + r - i, row[i]])
                    i = r
                else:
                    n = min(128, w - i)
                    body += bytes([n]) + row[i:i + n].tobytes()
                    i += n
    files = {
        "rle": b"#?RADIANCE\n# comment\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y 3  +X 200\n" + body,
        "rgbe_id": b"#?RGBE\nFORMAT=32-bit_rle_rgbe\n\n-Y 3 +X 200\n" + body,
        # the second scanline does not open with 2 2: stb reads the rest as
        # flat pixels starting over at the first pixel
        "fallback": b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 3 +X 200\n" + body[:first_line] +
                    pix.reshape(-1, 4).tobytes(),
        "flat": b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 3 +X 7\n" + pix[:, :7].tobytes(),
    }
    for name, data in files.items():
        path = str(tmp_path / (name + ".hdr"))
        open(path, "wb").write(data)
        ref = _ref_read(path, str(tmp_path))
        assert np.array_equal(g.image_read(path), ref), name
    # closed form: texel = c * 2^(e - 136), 0 for e = 0
    got = g.image_read(str(tmp_path / "rle.hdr"))
    e = pix[..., 3].astype(np.int64)
    expect = np.where((e > 0)[..., None], pix[..., :3] * np.ldexp(np.float32(1), e - 136)[..., None], 0).astype(np.float32)
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("text", [b"#?RADIANCE\n\n-Y 2 +X 2\n", b"P6\n2 2\n255\n",
                                  b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+X 2 -Y 2\n"])
def test_hdr_read_rejects_bad_files(tmp_path, text):
    path = str(tmp_path / "bad.hdr")
    open(path, "wb").write(text + bytes(64))
    with pytest.raises(g.GyError):
        g.image_read(path)
    with pytest.raises(g.GyError):
        g.image_read(str(tmp_path / "x.png"))

"""Image files (clv_write_png / clv_write_exr, clovers-cli/src/write.rs),
decoded here with zlib and struct: the PNG's pixels against the sRGB
encoding of the same Xyz<E> values computed from the colour module's
matrices, with closed forms (E white -> 255s, black -> 0s, an sRGB primary
comes back as itself); the EXR's floats bit for bit."""
import struct
import zlib

import numpy as np
import pytest

import clovers_amd
from clovers_amd.color import M_D65_TO_E, M_SRGB_TO_XYZ_D65, srgb_to_xyz_e


def png_chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, out = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        out.append((typ, body))
        pos += 12 + n
    return out


def read_png(path):
    ch = png_chunks(open(path, "rb").read())
    assert ch[0][0] == b"IHDR" and ch[-1][0] == b"IEND"
    W, H, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", ch[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in ch if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(H, 1 + 3 * W)
    assert np.all(rows[:, 0] == 0)
    texts = dict(b.split(b"\0", 1) for t, b in ch if t == b"tEXt")
    return rows[:, 1:].reshape(H, W, 3), texts


def expected_srgb8(xyz):
    M = np.linalg.inv(M_D65_TO_E @ M_SRGB_TO_XYZ_D65)
    lin = np.einsum("ij,...j->...i", M, xyz.astype(np.float64))
    lin = np.where(np.isnan(lin), 0.0, np.maximum(lin, 0))        # NaN mixes into every channel and encodes as 0
    enc = np.where(lin <= 0.0031308, 12.92 * lin, 1.055 * lin ** (1 / 2.4) - 0.055)
    return np.floor(np.clip(enc, 0, 1) * 255 + 0.5).astype(np.uint8)


def test_png_pixels_and_metadata(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1.2, size=(7, 11, 3)).astype(np.float32)
    img[0, 0] = 1.0                                   # E white
    img[0, 1] = 0.0
    img[0, 2] = srgb_to_xyz_e([1.0, 0.0, 0.0])       # the sRGB red primary
    img[0, 3] = 0.5
    img[0, 4] = [-1.0, np.nan, 0.2]
    p = tmp_path / "a.png"
    clovers_amd.write_png(img, p, comment="scene x.json, 64 spp")
    px, texts = read_png(p)
    assert px.shape == (7, 11, 3)
    assert list(px[0, 0]) == [255, 255, 255] and list(px[0, 1]) == [0, 0, 0]
    assert list(px[0, 2]) == [255, 0, 0]
    assert list(px[0, 3]) == [188] * 3                # encode(0.5) = 0.7354 -> 187.5 -> 188
    assert list(px[0, 4]) == [0, 0, 0]
    want = expected_srgb8(img)
    near = np.abs(px.astype(int) - want.astype(int))
    assert near.max() <= 1 and (near > 0).mean() < 0.01
    assert texts[b"Comment"] == b"scene x.json, 64 spp" and texts[b"Software"] == b"clovers_amd"


def read_exr(path):
    d = open(path, "rb").read()
    assert d[:8] == bytes([0x76, 0x2F, 0x31, 0x01, 2, 0, 0, 0])
    pos, attrs = 8, {}
    while d[pos] != 0:
        e = d.index(b"\0", pos)
        name = d[pos:e].decode()
        e2 = d.index(b"\0", e + 1)
        typ = d[e + 1:e2].decode()
        n, = struct.unpack("<i", d[e2 + 1:e2 + 5])
        attrs[name] = (typ, d[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack("<%dQ" % H, d[pos:pos + 8 * H])
    img = np.zeros((H, W, 3), np.float32)
    for y, o in enumerate(offs):
        yy, n = struct.unpack("<ii", d[o:o + 8])
        assert yy == y and n == 12 * W
        line = np.frombuffer(d[o + 8:o + 8 + n], "<f4").reshape(3, W)   # B, G, R
        img[y] = line[::-1].T
    return img, attrs


def test_exr_round_trip(tmp_path):
    img = np.random.default_rng(2).normal(size=(5, 9, 3)).astype(np.float32)
    p = tmp_path / "a.exr"
    clovers_amd.write_exr(img, p)
    back, attrs = read_exr(p)
    assert np.array_equal(back, img)
    assert attrs["compression"] == ("compression", b"\0")
    ch = np.frombuffer(attrs["chromaticities"][1], "<f4")
    assert np.allclose(ch, [1, 0, 0, 1, 0, 0, 1 / 3, 1 / 3])
    assert attrs["channels"][0] == "chlist"


def test_bad_arguments(tmp_path):
    with pytest.raises(clovers_amd.ClvError):
        clovers_amd.write_png(np.zeros((2, 2, 3)), tmp_path / "no" / "such" / "dir.png")
    with pytest.raises(ValueError):
        clovers_amd.write_exr(np.zeros((2, 2)), tmp_path / "x.exr")

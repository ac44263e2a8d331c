"""Film output: the 8-bit conversion of film::write (cameras/film.cpp:114-153,
utilities.cpp:33-38), the PNG and PFM writers.  Host code of librbw.so; no GPU.

The byte conversion is pinned by the sRGB standard's published values (IEC 61966-2-1:
linear 0.5 -> 188, 0.18 -> 118, the 0.0031308 knee), by the round trip through the
standard's inverse for all 256 codes, and by a float64 evaluation away from rounding
boundaries.  The PNG is decoded here from its chunks (CRC-32, zlib) and must give back
the same bytes; the PFM must give back the same floats."""
import math
import struct
import zlib

import numpy as np
import pytest

from rainbow_amd import abi
from rainbow_amd.tracer import RbwError, film_to_rgba8, write_pfm, write_png


def srgb64(v):
    """IEC 61966-2-1 encoding in float64"""
    return 12.92 * v if v <= 0.0031308 else 1.055 * v ** (1 / 2.4) - 0.055


def srgb_inverse(c):
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


def one_pixel(v, scale=1.0):
    img = np.full((1, 1, 3), v, dtype=np.float32)
    return film_to_rgba8(img, scale=scale)[0, 0]


def test_published_srgb_codes():
    assert one_pixel(0.0)[:3].tolist() == [0, 0, 0]
    assert one_pixel(1.0)[:3].tolist() == [255, 255, 255]
    assert one_pixel(0.5)[0] == 188
    assert one_pixel(0.18)[0] == 118
    # the linear segment up to the knee: 12.92 * 0.0031308 = 0.04045 -> 10.81 -> 10
    assert one_pixel(0.0031308)[0] == 10
    assert one_pixel(0.001)[0] == int(255 * 12.92 * 0.001 + 0.5)
    assert one_pixel(0.5)[3] == 255


def test_clamp_scale_and_special_values():
    assert one_pixel(-3.0)[0] == 0
    assert one_pixel(7.0)[0] == 255
    assert one_pixel(float("inf"))[0] == 255
    assert one_pixel(float("-inf"))[0] == 0
    assert one_pixel(float("nan"))[0] == 0
    assert one_pixel(0.25, scale=2.0)[0] == one_pixel(0.5)[0]


def test_round_trip_every_code():
    codes = np.arange(256)
    lin = np.array([srgb_inverse(c / 255.0) for c in codes], dtype=np.float32)
    img = np.repeat(lin[None, :, None], 3, axis=2)
    out = film_to_rgba8(img)
    assert out[0, :, 0].tolist() == codes.tolist()


def test_matches_float64_away_from_boundaries():
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.random(4000) * 1.2, rng.random(1000) * 0.004]).astype(np.float32)
    out = film_to_rgba8(np.repeat(v[None, :, None], 3, axis=2))[0, :, 0]
    x = np.array([255 * srgb64(float(a)) + 0.5 for a in v])
    frac = x - np.floor(x)
    far = (frac > 1e-4) & (frac < 1 - 1e-4)
    want = np.clip(np.floor(x), 0, 255).astype(np.uint8)
    assert far.sum() > 4900
    assert np.array_equal(out[far], want[far])


def test_pixels_outside_the_bound_are_white():
    img = np.zeros((4, 5, 3), dtype=np.float32)
    out = film_to_rgba8(img, bound_min=(1, 1), bound_max=(4, 3))
    inside = np.zeros((4, 5), dtype=bool)
    inside[1:3, 1:4] = True
    assert (out[inside][:, :3] == 0).all() and (out[inside][:, 3] == 255).all()
    assert (out[~inside] == 255).all()


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) == crc, kind
        chunks.append((kind, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for k, b in chunks if k == b"IDAT"))
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()  # filter type none
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    img = (rng.random((37, 53, 3)) * 1.3).astype(np.float32)
    p = tmp_path / "film.png"
    write_png(p, img, bound_min=(2, 3), bound_max=(50, 30), scale=0.8)
    assert np.array_equal(read_png(p), film_to_rgba8(img, bound_min=(2, 3), bound_max=(50, 30), scale=0.8))


def test_pfm_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.normal(size=(19, 23, 3)).astype(np.float32)
    img[0, 0, 0] = np.inf
    p = tmp_path / "film.pfm"
    write_pfm(p, img)
    data = open(p, "rb").read()
    header = b"PF\n23 19\n-1.0\n"
    assert data.startswith(header)
    back = np.frombuffer(data[len(header):], dtype="<f4").reshape(19, 23, 3)[::-1]
    assert np.array_equal(back, img)


def test_writer_errors(tmp_path):
    img = np.zeros((2, 2, 3), dtype=np.float32)
    with pytest.raises(RbwError) as e:
        write_png(tmp_path / "no_such_dir" / "x.png", img)
    assert e.value.code == abi.RBW_ERR_INVALID_ARGUMENT and "cannot open" in str(e.value)
    with pytest.raises(RbwError):
        write_pfm(tmp_path / "x.pfm", np.zeros((0, 2, 3), dtype=np.float32))

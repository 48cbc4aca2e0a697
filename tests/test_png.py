"""PNG decoding on the host (chunks, inflate, all five scanline filters, Adam7,
palettes, 16-bit samples, refusals), against Pillow's decoder on files
written by Pillow (adaptive filters) and by a writer of our own that takes
the specification literally (each filter type in turn, interlaced files).

The refusals follow the reference's tests and reader
(unit-tests/test_generic_readers.py:111-131, 261-271; generic_readers.py:
54-78): gray + alpha by name, colour types outside the specification, files
that are not PNG.  The gray conversion runs on the device
(tests/test_gpu_png.py)."""

import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from pivtools_amd import png as P

H, W = 16, 24
PNG_GRAY, PNG_RGB, PNG_PALETTE, PNG_GRAY_ALPHA, PNG_RGBA = 0, 2, 3, 4, 6
_SAMPLES = {PNG_GRAY: 1, PNG_RGB: 3, PNG_PALETTE: 1, PNG_GRAY_ALPHA: 2, PNG_RGBA: 4}


def _chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body))


def _filter_rows(rows: np.ndarray, bpp: int, start: int) -> bytes:
    """Rows of bytes filtered with types start, start+1, ... (mod 5)."""
    out = []
    prev = np.zeros(rows.shape[1], np.int32)
    for k, row in enumerate(rows.astype(np.int32)):
        t = (start + k) % 5
        left = np.concatenate([np.zeros(bpp, np.int32), row[:-bpp]]) if row.size > bpp else np.zeros_like(row)
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]]) if row.size > bpp else np.zeros_like(row)
        if t == 0:
            f = row
        elif t == 1:
            f = row - left
        elif t == 2:
            f = row - prev
        elif t == 3:
            f = row - (left + prev) // 2
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            f = row - pred
        out.append(bytes([t]) + (f & 0xFF).astype(np.uint8).tobytes())
        prev = row
    return b"".join(out)


def write_png(path, samples, colour_type, palette=None, trns=None, depth=None, interlace=False, filters=True):
    """A PNG from the specification.  samples: (H, W) or (H, W, spp) in the
    file's own order, uint8 / uint16, or palette indices / sub-byte gray
    values with depth given."""
    data = np.asarray(samples)
    data = data.reshape(data.shape[0], data.shape[1], -1)
    assert data.shape[2] == _SAMPLES[colour_type]
    depth = depth or (16 if data.dtype == np.uint16 else 8)
    h, w, ch = data.shape
    bpp = max(1, ch * depth // 8)

    def rows_of(img):
        if depth == 16:
            return img.astype(">u2").reshape(img.shape[0], -1).view(np.uint8)
        if depth == 8:
            return img.astype(np.uint8).reshape(img.shape[0], -1)
        bits = ((img[:, :, 0][:, :, None] >> np.arange(depth - 1, -1, -1)) & 1).astype(np.uint8)
        return np.packbits(bits.reshape(img.shape[0], -1), axis=1)

    raw = b""
    passes = P._ADAM7 if interlace else ((0, 0, 1, 1),)
    for k, (r0, c0, rs, cs) in enumerate(passes):
        sub = data[r0::rs, c0::cs]
        if sub.size == 0:
            continue
        rows = rows_of(sub)
        raw += _filter_rows(rows, bpp, k) if filters else b"".join(b"\x00" + r.tobytes() for r in rows)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, colour_type, 0, 0, 1 if interlace else 0)
    chunks = _chunk(b"IHDR", ihdr)
    if palette is not None:
        chunks += _chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        chunks += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw)
    chunks += _chunk(b"IDAT", z[:len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:]) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(P.SIGNATURE + chunks)
    return str(path)


def _ramp16():
    return np.linspace(0, 65535, H * W).reshape(H, W).astype(np.uint16)


# ---------------------------------------------------------------- values kept
def test_gray16_keeps_every_value(tmp_path):
    ramp = _ramp16()
    out = P.png_samples(write_png(tmp_path / "g.png", ramp, PNG_GRAY))
    assert out.dtype == np.uint16
    np.testing.assert_array_equal(out, ramp)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,dtype", [(PNG_GRAY, np.uint8), (PNG_GRAY, np.uint16), (PNG_RGB, np.uint8),
                                         (PNG_RGB, np.uint16), (PNG_RGBA, np.uint8), (PNG_RGBA, np.uint16)])
def test_every_filter_and_adam7_against_pillow(tmp_path, interlace, ctype, dtype):
    rng = np.random.default_rng(ctype * 7 + np.dtype(dtype).itemsize)
    shape = (37, 29) if ctype == PNG_GRAY else (37, 29, _SAMPLES[ctype])
    data = rng.integers(0, np.iinfo(dtype).max, shape, endpoint=True).astype(dtype)
    data[10:20] = data[10:11]                          # repeated rows for Up and Average
    path = write_png(tmp_path / "f.png", data, ctype, interlace=interlace)
    out = P.png_samples(path)
    np.testing.assert_array_equal(out, data)
    if dtype == np.uint8:
        np.testing.assert_array_equal(np.array(Image.open(path)), data)
    elif ctype == PNG_GRAY:
        np.testing.assert_array_equal(np.array(Image.open(path)).astype(np.uint16), data)


@pytest.mark.parametrize("mode", ["L", "RGB", "RGBA"])
def test_pillow_written_files(tmp_path, mode):
    rng = np.random.default_rng(11)
    shape = {"L": (61, 83), "RGB": (61, 83, 3), "RGBA": (61, 83, 4)}[mode]
    data = rng.integers(0, 256, shape).astype(np.uint8)
    data[:, 30:50] = 200                               # smooth areas: Pillow picks several filters
    path = str(tmp_path / "p.png")
    Image.fromarray(data, mode).save(path, optimize=True)
    np.testing.assert_array_equal(P.png_samples(path), data)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
@pytest.mark.parametrize("interlace", [False, True])
def test_palettes_expand_to_their_colours(tmp_path, depth, interlace):
    rng = np.random.default_rng(depth)
    n = 1 << depth
    pal = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    idx = rng.integers(0, n, (21, 35)).astype(np.uint8)
    path = write_png(tmp_path / "p.png", idx, PNG_PALETTE, palette=pal, depth=depth, interlace=interlace)
    out = P.png_samples(path)
    np.testing.assert_array_equal(out, pal[idx])
    np.testing.assert_array_equal(np.array(Image.open(path).convert("RGB")), pal[idx])
    trns = write_png(tmp_path / "t.png", idx, PNG_PALETTE, palette=pal, depth=depth, trns=[0, 128])
    out = P.png_samples(trns)
    assert out.shape == (21, 35, 4)
    np.testing.assert_array_equal(out[:, :, :3], pal[idx])


# ---------------------------------------------------------------- refusals
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16], ids=["la8", "la16"])
def test_gray_alpha_png_is_refused_by_name(tmp_path, dtype):
    path = write_png(tmp_path / "gray_alpha.png", np.zeros((H, W, 2), dtype=dtype), PNG_GRAY_ALPHA)
    with pytest.raises(ValueError, match=r"gray_alpha\.png.*gray\+alpha"):
        P.png_samples(path)


def test_header_refusals(tmp_path):
    bad = tmp_path / "bad.png"
    bad.write_bytes(b"GIF89a" + b"\0" * 40)
    with pytest.raises(ValueError, match=r"bad\.png does not start with a PNG signature"):
        P.png_samples(str(bad))
    p = write_png(tmp_path / "c.png", np.zeros((4, 4), np.uint8), PNG_GRAY)
    raw = bytearray(open(p, "rb").read())
    raw[25] = 5                                        # colour type byte of the IHDR
    (tmp_path / "c5.png").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match=r"c5\.png has PNG colour type 5"):
        P.png_samples(str(tmp_path / "c5.png"))
    g4 = write_png(tmp_path / "g4.png", np.zeros((4, 8), np.uint8), PNG_GRAY, depth=4)
    with pytest.raises(ValueError, match=r"4-bit gray"):
        P.png_samples(g4)


def test_damage_is_reported(tmp_path):
    p = write_png(tmp_path / "d.png", np.arange(64, dtype=np.uint8).reshape(8, 8), PNG_GRAY)
    raw = bytearray(open(p, "rb").read())
    i = raw.index(b"IDAT") + 6
    raw[i] ^= 0xFF
    (tmp_path / "crc.png").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match=r"IDAT fails its CRC"):
        P.png_samples(str(tmp_path / "crc.png"))
    (tmp_path / "cut.png").write_bytes(bytes(open(p, "rb").read()[:-12]))
    with pytest.raises(ValueError, match=r"truncated|past the end"):
        P.png_samples(str(tmp_path / "cut.png"))

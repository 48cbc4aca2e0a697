"""TIFF reading on the host (structure, strips, tiles, compressions,
predictor, refusals), checked against files written by Pillow's libtiff
(an independent encoder: LZW with 12-bit codes and clears, Deflate,
PackBits, horizontal differencing) and against a small writer of our own for
the layouts Pillow does not write (planar samples, tiles, big-endian,
BigTIFF, volumetric pages, signed and float samples).

The refusals and the kept values follow the reference's own tests
(unit-tests/test_generic_readers.py:37-84): multi-page files name their page
count, volumetric pages their axes 'ZYX', grayscale pages keep every column
and value.  The gray conversion itself runs on the device
(tests/test_gpu_tiff.py)."""

import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from pivtools_amd import _native
from pivtools_amd import tiff as T

H, W = 16, 24

# ---------------------------------------------------------------- writer
_TYPES = {np.dtype(np.uint8): (1, 8), np.dtype(np.uint16): (1, 16), np.dtype(np.int16): (2, 16),
          np.dtype(np.int8): (2, 8), np.dtype(np.uint32): (1, 32), np.dtype(np.int32): (2, 32),
          np.dtype(np.float32): (3, 32)}


def write_tiff(path, data, *, planar=False, tile=None, bo="<", big=False, compression=1, predictor=1,
               depth=None, extra_tags=(), pages=1, rows_per_strip=None, photometric=None, count_extra=0):
    """A TIFF from TIFF 6.0 itself.  data: (H, W), (H, W, S) (contiguous) or
    (S, H, W) with planar=True; depth: an ImageDepth tag (volumetric).
    compression 1 (raw) or 8 (zlib); predictor 2 differences each row."""
    data = np.asarray(data)
    if planar:
        S, h, w = data.shape
        planes = [data[s][:, :, None] for s in range(S)]
    else:
        h, w = data.shape[:2]
        S = 1 if data.ndim == 2 else data.shape[2]
        planes = [data.reshape(h, w, S)]
    fmt, bits = _TYPES[data.dtype]
    dt = data.dtype.newbyteorder(bo)

    def encode(block):
        b = np.ascontiguousarray(block)
        if predictor == 2:
            d = b.copy()
            d[:, 1:] = b[:, 1:] - b[:, :-1]
            b = d
        raw = b.astype(dt).tobytes()
        return zlib.compress(raw) if compression == 8 else raw

    blobs = []
    for p in planes:
        if tile is None:
            rps = rows_per_strip or h
            for r0 in range(0, h, rps):
                blobs.append(encode(p[r0:r0 + rps]))
        else:
            tl, tw = tile
            for r0 in range(0, h, tl):
                for c0 in range(0, w, tw):
                    t = np.zeros((tl, tw, p.shape[2]), p.dtype)
                    blk = p[r0:r0 + tl, c0:c0 + tw]
                    t[:blk.shape[0], :blk.shape[1]] = blk
                    blobs.append(encode(t))
    off_t, cnt_t = (324, 325) if tile is not None else (273, 279)
    head = 16 if big else 8
    data_off = head
    offsets = []
    for b in blobs:
        offsets.append(data_off)
        data_off += len(b)
    if photometric is None:
        photometric = 2 if S in (3, 4) else 1
    tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bits] * S), (259, 3, [compression]), (262, 3, [photometric]),
            (277, 3, [S]), (284, 3, [2 if planar else 1]), (317, 3, [predictor]), (339, 3, [fmt] * S),
            (off_t, 16 if big else 4, offsets),
            (cnt_t, 16 if big else 4, [len(b) + (count_extra if k == 0 else 0) for k, b in enumerate(blobs)])]
    if tile is None:
        tags.append((278, 4, [rows_per_strip or h]))
    else:
        tags += [(322, 4, [tile[1]]), (323, 4, [tile[0]])]
    if depth is not None:
        tags.append((32997, 4, [depth]))
    tags += list(extra_tags)
    tags.sort()
    codes = {3: "H", 4: "I", 16: "Q"}
    size = {3: 2, 4: 4, 16: 8}
    ifds = []
    ifd_off = data_off + (data_off & 1)
    vsize, esize = (8, 20) if big else (4, 12)
    for pg in range(pages):
        n = len(tags)
        ifd_len = (8 if big else 2) + n * esize + vsize
        ext_off = ifd_off + ifd_len
        body, ext = b"", b""
        for tag, typ, vals in tags:
            raw = struct.pack(bo + codes[typ] * len(vals), *vals)
            cnt = struct.pack(bo + ("Q" if big else "I"), len(vals))
            if len(raw) <= vsize:
                val = raw + b"\0" * (vsize - len(raw))
            else:
                val = struct.pack(bo + ("Q" if big else "I"), ext_off + len(ext))
                ext += raw
            body += struct.pack(bo + "HH", tag, typ) + cnt + val
        nxt = ifd_off + ifd_len + len(ext) if pg + 1 < pages else 0
        nxt += nxt & 1
        ifd = struct.pack(bo + ("Q" if big else "H"), n) + body + struct.pack(bo + ("Q" if big else "I"), nxt)
        blob = ifd + ext
        blob += b"\0" * (len(blob) & 1)
        ifds.append(blob)
        ifd_off = nxt
    first = data_off + (data_off & 1)
    hdr = (b"II" if bo == "<" else b"MM") + (struct.pack(bo + "HHHQ", 43, 8, 0, first) if big
                                             else struct.pack(bo + "HI", 42, first))
    with open(path, "wb") as f:
        f.write(hdr + b"".join(blobs) + b"\0" * (data_off & 1) + b"".join(ifds))
    return str(path)


def _samples(path):
    return T.tiff_samples(path)[0]


# ---------------------------------------------------------------- reference's cases
@pytest.mark.parametrize(("pages", "width"), [(2, W), (5, W), (2, 3), (2, 4)], ids=str)
def test_multipage_tiff_is_refused_with_its_page_count(tmp_path, pages, width):
    data = np.arange(pages * H * width, dtype=np.uint16).reshape(pages, H, width)
    path = str(tmp_path / "stack.tif")
    Image.fromarray(data[0]).save(path, save_all=True, append_images=[Image.fromarray(d) for d in data[1:]])
    assert T.tiff_page_count(path) == pages
    with pytest.raises(ValueError, match=rf"stack\.tif.*{pages} pages"):
        T.tiff_samples(path)
    # our writer's chain too
    p2 = write_tiff(tmp_path / "stack2.tif", data[0], pages=pages)
    with pytest.raises(ValueError, match=rf"stack2\.tif.*{pages} pages"):
        T.tiff_samples(p2)


def test_volumetric_tiff_page_is_refused_with_its_axes(tmp_path):
    path = write_tiff(tmp_path / "volume.tif", np.zeros((H, H), np.uint16), tile=(H, H), depth=4)
    with pytest.raises(ValueError, match=r"volume\.tif.*axes 'ZYX'"):
        T.tiff_samples(path)


@pytest.mark.parametrize("width", [3, 4, W])
def test_grayscale_tiff_keeps_every_column_and_value(tmp_path, width):
    data = np.arange(H * width, dtype=np.uint16).reshape(H, width) * 97
    path = str(tmp_path / "gray.tif")
    Image.fromarray(data).save(path)
    out, axes, code = T.tiff_samples(path)
    assert axes == "YX" and code == 1
    np.testing.assert_array_equal(out, data)


# ---------------------------------------------------------------- Pillow / libtiff files
@pytest.mark.parametrize("comp", ["raw", "tiff_lzw", "tiff_deflate", "tiff_adobe_deflate", "packbits"])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("predictor", [1, 2])
def test_pillow_written_gray_pages(tmp_path, comp, dtype, predictor):
    if predictor == 2 and comp in ("raw", "packbits"):
        pytest.skip("libtiff applies the predictor to LZW and Deflate only")
    rng = np.random.default_rng(3)
    data = rng.integers(0, np.iinfo(dtype).max, (61, 77)).astype(dtype)
    data[20:40] = np.arange(77, dtype=dtype)          # runs and ramps for PackBits and the predictor
    path = str(tmp_path / "p.tif")
    info = {317: predictor} if predictor == 2 else {}
    Image.fromarray(data).save(path, compression=comp, tiffinfo=info)
    pages, page = T.parse_tiff(path)
    assert pages == 1 and page.predictor == predictor
    np.testing.assert_array_equal(_samples(path), data)


def test_lzw_with_wide_codes_and_clears(tmp_path):
    """512 x 512 noise: the libtiff encoder runs the table to 12-bit codes and
    emits clears; strips of 16 rows each restart the table."""
    rng = np.random.default_rng(0)
    data = rng.integers(0, 256, (512, 512), np.uint8)
    data[100:300, 50:400] = 7
    path = str(tmp_path / "n.tif")
    Image.fromarray(data).save(path, compression="tiff_lzw")
    _, page = T.parse_tiff(path)
    assert page.compression == 5 and len(page.offsets) > 1
    np.testing.assert_array_equal(_samples(path), data)


@pytest.mark.parametrize("comp", ["raw", "tiff_lzw", "tiff_deflate"])
def test_pillow_written_rgb8_contiguous(tmp_path, comp):
    rgb = np.random.default_rng(1).integers(0, 256, (33, 45, 3), np.uint8)
    path = str(tmp_path / "rgb.tif")
    Image.fromarray(rgb).save(path, compression=comp)
    out, axes, code = T.tiff_samples(path)
    assert axes == "YXS" and code == 0
    np.testing.assert_array_equal(out, rgb)


# ---------------------------------------------------------------- our writer's layouts
@pytest.mark.parametrize("compression", [1, 8])
def test_planar_rgb_reads_as_sample_planes(tmp_path, compression):
    rgb = np.random.default_rng(2).integers(0, 65535, (3, 20, 30)).astype(np.uint16)
    path = write_tiff(tmp_path / "pl.tif", rgb, planar=True, compression=compression, rows_per_strip=7)
    out, axes, _ = T.tiff_samples(path)
    assert axes == "SYX"
    np.testing.assert_array_equal(out, rgb)


@pytest.mark.parametrize("predictor", [1, 2])
@pytest.mark.parametrize("compression", [1, 8])
def test_tiles_with_ragged_edges(tmp_path, predictor, compression):
    data = np.random.default_rng(4).integers(0, 65535, (37, 53)).astype(np.uint16)
    path = write_tiff(tmp_path / "t.tif", data, tile=(16, 16), predictor=predictor, compression=compression)
    _, page = T.parse_tiff(path)
    assert page.tile == (16, 16) and len(page.offsets) == 3 * 4
    # libtiff (and so this reader) applies a predictor to compressed data only:
    # uncompressed tiles hold their row differences as stored
    want = data
    if predictor == 2 and compression == 1:
        want = np.empty_like(data)
        for r in range(0, 37, 16):
            for c in range(0, 53, 16):
                t = np.zeros((16, 16), data.dtype)
                blk = data[r:r + 16, c:c + 16]
                t[:blk.shape[0], :blk.shape[1]] = blk
                t[:, 1:] = t[:, 1:] - t[:, :-1].copy()
                want[r:r + 16, c:c + 16] = t[:blk.shape[0], :blk.shape[1]]
    np.testing.assert_array_equal(_samples(path), want)
    # the independent decoder reads the same file the same way
    np.testing.assert_array_equal(np.array(Image.open(path)), want)


@pytest.mark.parametrize("bo", ["<", ">"])
@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("dtype", [np.uint16, np.int16, np.float32, np.int32])
def test_byte_orders_bigtiff_and_sample_types(tmp_path, bo, big, dtype):
    rng = np.random.default_rng(5)
    data = (rng.normal(0, 1000, (19, 23))).astype(dtype)
    path = write_tiff(tmp_path / "b.tif", data, bo=bo, big=big, rows_per_strip=5, compression=8,
                      predictor=2 if dtype != np.float32 else 1)
    out, _, code = T.tiff_samples(path)
    assert out.dtype == np.dtype(dtype) and out.dtype.isnative
    np.testing.assert_array_equal(out, data)
    assert code == T.SAMPLE_TYPES[{np.uint16: (1, 16), np.int16: (2, 16), np.float32: (3, 32),
                                   np.int32: (2, 32)}[dtype]][1]


def test_two_samples_and_rgba(tmp_path):
    ga = np.random.default_rng(6).integers(0, 255, (12, 14, 2)).astype(np.uint8)
    out, axes, _ = T.tiff_samples(write_tiff(tmp_path / "ga.tif", ga, extra_tags=[(338, 3, [2])]))
    assert axes == "YXS" and out.shape == (12, 14, 2)
    rgba = np.random.default_rng(7).integers(0, 255, (12, 14, 4)).astype(np.uint8)
    out, axes, _ = T.tiff_samples(write_tiff(tmp_path / "rgba.tif", rgba, extra_tags=[(338, 3, [2])]))
    np.testing.assert_array_equal(out, rgba)


def test_refusals_name_what_they_refuse(tmp_path):
    d = np.zeros((8, 8), np.uint16)
    with pytest.raises(ValueError, match=r"compression 7 \(JPEG\)"):
        T.tiff_samples(write_tiff(tmp_path / "j.tif", d, compression=7))
    with pytest.raises(ValueError, match=r"predictor 3"):
        T.tiff_samples(write_tiff(tmp_path / "p3.tif", d, predictor=3))
    rgb16s = np.zeros((8, 8, 3), np.int16)
    with pytest.raises(ValueError, match=r"colour samples of type int16"):
        T.tiff_samples(write_tiff(tmp_path / "s.tif", rgb16s))
    (tmp_path / "x.tif").write_bytes(b"not a tiff at all")
    with pytest.raises(ValueError, match=r"byte-order mark"):
        T.tiff_samples(str(tmp_path / "x.tif"))
    with pytest.raises(ValueError, match=r"block 0 runs past the end of the file"):
        T.tiff_samples(write_tiff(tmp_path / "cut.tif", d, count_extra=1 << 20))
    with pytest.raises(ValueError, match=r"decoded 100 bytes, expected 128"):
        T.tiff_samples(write_tiff(tmp_path / "short.tif", d, count_extra=-28))


# ---------------------------------------------------------------- the serial decoders
def test_packbits_known_answer():
    """TIFF 6.0 section 9's worked example."""
    L = _native.core()
    src = np.frombuffer(bytes([0xFE, 0xAA, 0x02, 0x80, 0x00, 0x2A, 0xFD, 0xAA, 0x03, 0x80, 0x00, 0x2A, 0x22,
                               0xF7, 0xAA]), np.uint8)
    want = bytes([0xAA] * 3 + [0x80, 0x00, 0x2A] + [0xAA] * 4 + [0x80, 0x00, 0x2A, 0x22] + [0xAA] * 10)
    dst = np.zeros(64, np.uint8)
    n = L.pivamd_packbits(src.ctypes.data, src.size, dst.ctypes.data, 64)
    assert dst[:n].tobytes() == want
    n = L.pivamd_packbits(src.ctypes.data, src.size, dst.ctypes.data, 5)
    assert n == 5 and dst[:5].tobytes() == want[:5]


def test_lzw_refuses_old_style_and_bad_codes():
    L = _native.core()
    dst = np.zeros(16, np.uint8)
    old = np.frombuffer(b"\x00\x01\x02\x03", np.uint8)
    assert L.pivamd_tiff_lzw(old.ctypes.data, old.size, dst.ctypes.data, 16) == -1
    # clear (256), then code 300 before the table holds it: 9-bit codes MSB first
    bits = format(256, "09b") + format(300, "09b")
    bits += "0" * (-len(bits) % 8)
    bad = np.frombuffer(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)), np.uint8)
    assert L.pivamd_tiff_lzw(bad.ctypes.data, bad.size, dst.ctypes.data, 16) == -1


def test_gray_conversion_refuses_bad_arguments_before_any_launch():
    """Argument checks of pivamd_gray_from_samples return before touching a
    device, so they run here: luma of a type OpenCV does not convert, and
    empty or oversized batches."""
    L = _native.core()
    p = 16                                              # any non-null address; nothing is read
    assert L.pivamd_gray_from_samples(p, 1, 4, 4, 3, 0, 3, p, None) == 1       # int16 RGB
    assert L.pivamd_gray_from_samples(p, 1, 4, 4, 4, 1, 5, p, None) == 1       # int32 RGBA
    assert L.pivamd_gray_from_samples(p, 1, 0, 4, 1, 0, 0, p, None) == 1       # empty frame
    assert L.pivamd_gray_from_samples(p, 70000, 4, 4, 1, 0, 0, p, None) == 1   # past the grid's y limit
    assert L.pivamd_gray_from_samples(p, 1, 4, 4, 1, 0, 9, p, None) == 1       # unknown sample type
    assert L.pivamd_gray_from_samples(p, 0, 4, 4, 1, 0, 0, p, None) == 0       # nothing to do

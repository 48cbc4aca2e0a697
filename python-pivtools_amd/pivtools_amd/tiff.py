"""TIFF frames: one page per file, decoded on the host, made gray on the device.

Follows the reference's ``read_tiff``
(``pivtools_core/image_handling/readers/generic_readers.py:81-122``), which
reads through tifffile and refuses what PIVTOOLs does not treat as one frame:

* a file with more than one page (``is a multi-page TIFF with N pages``);
* a page with an axis other than Y, X and S (a volumetric page: ``ZYX``).

A page's samples keep their stored type; the sample axis ``S`` exists when
SamplesPerPixel > 1 (last for contiguous, first for planar pages, as
tifffile orders them).  3 or 4 samples become BT.601 luma by OpenCV's
RGB2GRAY rule (``_luma``, ``:33-51``; a fourth sample is alpha), any other
count their mean, and the frame is float32.

Host work here: the TIFF structure (classic and BigTIFF, either byte order),
strips or tiles, and their compression (none; LZW and PackBits in the
library's host code, ``csrc/readers.cpp``; Deflate through zlib) and the
horizontal-differencing predictor.  The per-pixel conversion to float32
gray runs on the device (``pivamd_gray_from_samples``), for one frame or a
batch of same-shaped frames at once.  A layout this reader does not decode
(JPEG and other compressions, bit depths other than 8, 16 and 32, the
floating-point predictor, bit-reversed fill order) is refused by name.
"""

from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native

# TIFF 6.0 tags (and ImageDepth from SGI's extension, as tifffile reads it)
_T_WIDTH, _T_LENGTH, _T_BITS, _T_COMPRESSION, _T_PHOTOMETRIC = 256, 257, 258, 259, 262
_T_FILL_ORDER, _T_STRIP_OFFSETS, _T_SPP, _T_ROWS_PER_STRIP, _T_STRIP_COUNTS = 266, 273, 277, 278, 279
_T_PLANAR, _T_PREDICTOR, _T_TILE_W, _T_TILE_L, _T_TILE_OFFSETS, _T_TILE_COUNTS = 284, 317, 322, 323, 324, 325
_T_SAMPLE_FORMAT, _T_IMAGE_DEPTH = 339, 32997

COMPRESSIONS = {1: "none", 5: "LZW", 8: "Deflate", 32946: "Deflate", 32773: "PackBits"}
_REFUSED_COMPRESSIONS = {2: "CCITT RLE", 3: "CCITT G3", 4: "CCITT G4", 6: "old JPEG", 7: "JPEG",
                         34712: "JPEG 2000", 50000: "Zstandard", 34925: "LZMA", 50001: "WebP"}
# (SampleFormat, bits) -> numpy type and pivamd_gray_from_samples code
SAMPLE_TYPES = {(1, 8): (np.uint8, 0), (1, 16): (np.uint16, 1), (2, 8): (np.int8, 2), (2, 16): (np.int16, 3),
                (1, 32): (np.uint32, 4), (2, 32): (np.int32, 5), (3, 32): (np.float32, 6)}
_LUMA_TYPES = (0, 1, 6)      # uint8, uint16, float32: the types OpenCV's conversion takes
# the compressions a Predictor tag applies to (libtiff's LZW and Deflate codecs;
# uncompressed and PackBits data are stored as they are)
_PREDICTED = (5, 8, 32946)

# field type -> (struct code, bytes)
_FIELD = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1), 7: ("B", 1),
          8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8), 17: ("q", 8),
          18: ("Q", 8)}
_MAX_PAGES = 1 << 20


@dataclass
class TiffPage:
    """The first page's structure: everything the decode needs."""

    path: str
    byteorder: str              # '<' or '>'
    width: int
    height: int
    depth: int                  # ImageDepth (1: a 2-D page)
    spp: int
    bits: int
    sample_format: int          # 1 uint, 2 int, 3 float
    planar: int                 # 1 contiguous, 2 separate planes
    compression: int
    predictor: int
    photometric: int
    fill_order: int
    offsets: Tuple[int, ...]
    counts: Tuple[int, ...]
    rows_per_strip: int
    tile: Optional[Tuple[int, int]]    # (length, width) of a tile, None for strips

    @property
    def axes(self) -> str:
        """tifffile's axes of the page: Y X, with S for samples (first when
        planar) and Z for a volumetric page."""
        core = ("Z" if self.depth > 1 else "") + "YX"
        if self.spp > 1:
            return "S" + core if self.planar == 2 else core + "S"
        return core

    @property
    def sample_type(self) -> Tuple[type, int]:
        key = (self.sample_format, self.bits)
        if key not in SAMPLE_TYPES:
            raise ValueError(f"{Path(self.path).name} has {self.bits}-bit samples of TIFF SampleFormat "
                             f"{self.sample_format}; this reader decodes 8-, 16- and 32-bit integer samples "
                             "and 32-bit float samples")
        return SAMPLE_TYPES[key]


def _read_tags(buf: bytes, off: int, bo: str, big: bool, name: str) -> Tuple[dict, int]:
    """The tags of the IFD at off, and the next IFD's offset."""
    if big:
        if off + 8 > len(buf):
            raise ValueError(f"{name}: IFD offset {off} past the end of the file")
        (n,) = struct.unpack_from(bo + "Q", buf, off)
        ent, esize, vsize, base = off + 8, 20, 8, "Q"
    else:
        if off + 2 > len(buf):
            raise ValueError(f"{name}: IFD offset {off} past the end of the file")
        (n,) = struct.unpack_from(bo + "H", buf, off)
        ent, esize, vsize, base = off + 2, 12, 4, "I"
    end = ent + n * esize
    if end + vsize > len(buf):
        raise ValueError(f"{name}: IFD of {n} entries runs past the end of the file")
    tags = {}
    for k in range(n):
        e = ent + k * esize
        tag, typ = struct.unpack_from(bo + "HH", buf, e)
        (count,) = struct.unpack_from(bo + base, buf, e + 4)
        if typ not in _FIELD:
            continue
        code, size = _FIELD[typ]
        nbytes = size * count
        vpos = e + 4 + (8 if big else 4)
        if nbytes > vsize:
            (vpos,) = struct.unpack_from(bo + base, buf, vpos)
            if vpos + nbytes > len(buf):
                raise ValueError(f"{name}: tag {tag} values run past the end of the file")
        if typ == 2:
            tags[tag] = buf[vpos:vpos + count]
        elif typ in (5, 10):
            v = struct.unpack_from(bo + code[0] * (2 * count), buf, vpos)
            tags[tag] = tuple(v[i] / v[i + 1] if v[i + 1] else float("nan") for i in range(0, 2 * count, 2))
        else:
            tags[tag] = struct.unpack_from(bo + code * count, buf, vpos)
    (nxt,) = struct.unpack_from(bo + base, buf, end)
    return tags, nxt


def parse_tiff(path, buf: Optional[bytes] = None) -> Tuple[int, TiffPage]:
    """(number of pages, first page) of a TIFF file (classic or BigTIFF);
    buf: the file's bytes when the caller has read them."""
    path = str(path)
    name = Path(path).name
    if buf is None:
        buf = Path(path).read_bytes()
    if len(buf) < 8 or buf[:2] not in (b"II", b"MM"):
        raise ValueError(f"{name} does not start with a TIFF byte-order mark")
    bo = "<" if buf[:2] == b"II" else ">"
    (magic,) = struct.unpack_from(bo + "H", buf, 2)
    if magic == 42:
        big = False
        (first,) = struct.unpack_from(bo + "I", buf, 4)
    elif magic == 43:
        big = True
        if len(buf) < 16:
            raise ValueError(f"{name}: truncated BigTIFF header")
        bsize, _, first = struct.unpack_from(bo + "HHQ", buf, 4)
        if bsize != 8:
            raise ValueError(f"{name}: BigTIFF offset size {bsize}, expected 8")
    else:
        raise ValueError(f"{name} has TIFF version {magic}, expected 42 (TIFF) or 43 (BigTIFF)")
    pages, seen, off, tags0 = 0, set(), first, None
    while off and pages < _MAX_PAGES:
        if off in seen:
            raise ValueError(f"{name}: the IFD chain loops back to offset {off}")
        seen.add(off)
        tags, off = _read_tags(buf, off, bo, big, name)
        if tags0 is None:
            tags0 = tags
        pages += 1
    if tags0 is None:
        raise ValueError(f"{name} holds no image (empty IFD chain)")

    def one(tag, default=None):
        v = tags0.get(tag)
        if not v:
            if default is None:
                raise ValueError(f"{name}: required TIFF tag {tag} missing")
            return default
        return int(v[0])

    spp = one(_T_SPP, 1)
    bits = tags0.get(_T_BITS) or (1,)
    if len(set(bits)) != 1:
        raise ValueError(f"{name} has samples of different bit depths {tuple(bits)}")
    fmt = tags0.get(_T_SAMPLE_FORMAT) or (1,)
    if len(set(fmt)) != 1:
        raise ValueError(f"{name} has samples of different formats {tuple(fmt)}")
    tiled = _T_TILE_OFFSETS in tags0
    offs = tags0.get(_T_TILE_OFFSETS if tiled else _T_STRIP_OFFSETS)
    cnts = tags0.get(_T_TILE_COUNTS if tiled else _T_STRIP_COUNTS)
    if offs is None or cnts is None or len(offs) != len(cnts):
        raise ValueError(f"{name}: missing or mismatched {'tile' if tiled else 'strip'} offsets and byte counts")
    height = one(_T_LENGTH)
    if tiled and (one(_T_TILE_L) <= 0 or one(_T_TILE_W) <= 0):
        raise ValueError(f"{name}: tile size {one(_T_TILE_L)} x {one(_T_TILE_W)}")
    if not tiled and one(_T_ROWS_PER_STRIP, height) <= 0:
        raise ValueError(f"{name}: RowsPerStrip {one(_T_ROWS_PER_STRIP, height)}")
    page = TiffPage(path=path, byteorder=bo, width=one(_T_WIDTH), height=height, depth=one(_T_IMAGE_DEPTH, 1),
                    spp=spp, bits=int(bits[0]), sample_format=int(fmt[0]), planar=one(_T_PLANAR, 1),
                    compression=one(_T_COMPRESSION, 1), predictor=one(_T_PREDICTOR, 1),
                    photometric=one(_T_PHOTOMETRIC, 1), fill_order=one(_T_FILL_ORDER, 1),
                    offsets=tuple(int(x) for x in offs), counts=tuple(int(x) for x in cnts),
                    rows_per_strip=min(one(_T_ROWS_PER_STRIP, height), height) if not tiled else 0,
                    tile=(one(_T_TILE_L), one(_T_TILE_W)) if tiled else None)
    return pages, page


def _check_page(pages: int, page: TiffPage) -> None:
    name = Path(page.path).name
    if pages != 1:
        raise ValueError(f"{name} is a multi-page TIFF with {pages} pages. PIVTOOLs reads one image per TIFF "
                         "file and does not treat pages as frames. Save each image as its own TIFF.")
    if set(page.axes) - {"Y", "X", "S"}:
        raise ValueError(f"{name} has TIFF page axes {page.axes!r}. PIVTOOLs reads a 2-D page, optionally "
                         "with colour samples (axes YX, YXS or SYX).")
    if page.compression not in COMPRESSIONS:
        what = _REFUSED_COMPRESSIONS.get(page.compression, "unknown")
        raise ValueError(f"{name} uses TIFF compression {page.compression} ({what}), which this reader does "
                         f"not decode ({', '.join(sorted(set(COMPRESSIONS.values())))} are)")
    if page.predictor not in (1, 2):
        raise ValueError(f"{name} uses TIFF predictor {page.predictor}; this reader undoes 1 (none) and 2 "
                         "(horizontal differencing)")
    if page.fill_order != 1:
        raise ValueError(f"{name} has bit-reversed fill order (FillOrder {page.fill_order})")
    if page.planar not in (1, 2):
        raise ValueError(f"{name} has PlanarConfiguration {page.planar}")
    typ, code = page.sample_type
    if page.spp in (3, 4) and code not in _LUMA_TYPES:
        raise ValueError(f"{name}: colour samples of type {np.dtype(typ).name}; the BT.601 conversion takes "
                         "uint8, uint16 or float32 samples")
    if page.predictor == 2 and code == 6:
        raise ValueError(f"{name}: horizontal differencing on float samples")


def _decompress(raw: bytes, page: TiffPage, expected: int, where: str) -> bytes:
    c = page.compression
    if c == 1:
        out = raw[:expected]
    elif c in (8, 32946):
        try:
            out = zlib.decompress(raw)
        except zlib.error as e:
            raise ValueError(f"{where}: Deflate data damaged ({e})") from None
    else:
        L = _native.core()
        src = np.frombuffer(raw, np.uint8)
        dst = np.empty(max(expected, 1), np.uint8)
        fn = L.pivamd_tiff_lzw if c == 5 else L.pivamd_packbits
        got = int(fn(src.ctypes.data, src.size, dst.ctypes.data, expected))
        if got < 0:
            raise ValueError(f"{where}: {COMPRESSIONS[c]} data damaged or in the old bit-reversed form")
        out = dst[:got].tobytes()
    if len(out) < expected:
        raise ValueError(f"{where}: decoded {len(out)} bytes, expected {expected}")
    return out[:expected]


def tiff_samples(path) -> Tuple[np.ndarray, str, int]:
    """The first page's samples in native byte order, (H, W), (H, W, S) or
    (S, H, W) as its axes say, with the axes and the device sample code.
    Refuses what ``read_tiff`` refuses."""
    buf = Path(str(path)).read_bytes()
    pages, page = parse_tiff(path, buf)
    _check_page(pages, page)
    name = Path(page.path).name
    typ, code = page.sample_type
    dt = np.dtype(typ).newbyteorder(page.byteorder)
    H, W, S = page.height, page.width, page.spp
    sp = S if page.planar == 1 else 1          # samples per stored pixel
    planes = 1 if page.planar == 1 else S
    out = np.empty((planes, H, W, sp), dtype=np.dtype(typ))

    def block(k: int, rows: int, cols: int) -> np.ndarray:
        o, n = page.offsets[k], page.counts[k]
        if o + n > len(buf):
            raise ValueError(f"{name}: block {k} runs past the end of the file")
        expected = rows * cols * sp * dt.itemsize
        data = _decompress(buf[o:o + n], page, expected, f"{name} block {k}")
        a = np.frombuffer(data, dtype=dt).reshape(rows, cols, sp).astype(np.dtype(typ))
        if page.predictor == 2 and page.compression in _PREDICTED:
            # horizontal differencing restarts on every row of the block
            a = np.cumsum(a, axis=1, dtype=np.dtype(typ))
        return a

    if page.tile is None:
        rps = page.rows_per_strip
        per_plane = -(-H // rps)
        if len(page.offsets) < per_plane * planes:
            raise ValueError(f"{name}: {len(page.offsets)} strips, expected {per_plane * planes}")
        for p in range(planes):
            for s in range(per_plane):
                r0 = s * rps
                rows = min(rps, H - r0)
                out[p, r0:r0 + rows] = block(p * per_plane + s, rows, W)
    else:
        tl, tw = page.tile
        down, across = -(-H // tl), -(-W // tw)
        if len(page.offsets) < down * across * planes:
            raise ValueError(f"{name}: {len(page.offsets)} tiles, expected {down * across * planes}")
        for p in range(planes):
            for ty in range(down):
                for tx in range(across):
                    t = block((p * down + ty) * across + tx, tl, tw)
                    r0, c0 = ty * tl, tx * tw
                    out[p, r0:r0 + tl, c0:c0 + tw] = t[:H - r0, :W - c0]
    if S == 1:
        return out[0, :, :, 0], page.axes, code
    if page.planar == 2:
        return out[:, :, :, 0], page.axes, code
    return out[0], page.axes, code


def read_tiff(path, device=None):
    """One TIFF page as a float32 (H, W) device tensor."""
    return read_tiff_frames([path], device=device)[0]


def read_tiff_frames(paths: Sequence, device=None, stream_handle=None):
    """Same-shaped single-page TIFFs as one float32 (n, H, W) device tensor:
    the samples of all files cross PCIe once and one launch makes them gray."""
    import torch

    paths = list(paths)
    if not paths:
        raise ValueError("no TIFF files given")
    dev = torch.device(device) if device is not None else torch.device("cuda")
    frames, key = [], None
    for p in paths:
        a, axes, code = tiff_samples(p)
        k = (a.shape, axes, code)
        if key is None:
            key = k
        elif k != key:
            raise ValueError(f"{Path(str(p)).name}: samples {a.shape} {axes} of type {code} differ from the "
                             f"first file's {key[0]} {key[1]} of type {key[2]}")
        frames.append(a)
    shape, axes, code = key
    planar = 1 if axes.startswith("S") else 0
    S = shape[0] if planar else (shape[2] if len(shape) == 3 else 1)
    H, W = (shape[1], shape[2]) if planar else (shape[0], shape[1])
    host = np.ascontiguousarray(np.stack(frames))
    src = torch.from_numpy(host.view(np.uint8).reshape(-1)).to(dev)
    out = torch.empty((len(paths), H, W), dtype=torch.float32, device=dev)
    L = _native.core()
    st = stream_handle if stream_handle is not None else torch.cuda.current_stream(dev).cuda_stream
    _native.check(L.pivamd_gray_from_samples(src.data_ptr(), len(paths), H, W, S, planar, code, out.data_ptr(), st),
                  "gray_from_samples")
    return out


def tiff_page_count(path) -> int:
    return parse_tiff(path)[0]


def tiff_shape(path) -> Tuple[int, int]:
    _, page = parse_tiff(path)
    return page.height, page.width


__all__: List[str] = ["TiffPage", "parse_tiff", "tiff_samples", "read_tiff", "read_tiff_frames", "tiff_page_count",
                      "tiff_shape", "COMPRESSIONS", "SAMPLE_TYPES"]

"""PNG frames: decoded on the host, made gray on the device.

Follows the PNG half of the reference's ``read_png_jpeg``
(``pivtools_core/image_handling/readers/generic_readers.py:54-78, 125-176``),
which decodes through OpenCV at the file's own depth:

* gray (colour type 0) at 8 or 16 bits keeps every value;
* RGB, RGBA and palette files (types 2, 6, 3) become BT.601 luma by
  OpenCV's RGB2GRAY rule, a palette through its colours, alpha ignored;
* gray + alpha (type 4) is refused by name, and so is a colour type the
  specification does not define, read from the IHDR before decoding.

Host work: the chunk stream (CRC-checked), inflating the image data with
zlib, Adam7 interlacing and the scanline filters (``pivamd_png_unfilter``
in the library's host code), palette expansion.  The conversion to float32
gray is the device's ``pivamd_gray_from_samples``, shared with the TIFF
reader.  Gray files below 8 bits are refused (how OpenCV scales them is not
pinned here); JPEG is not decoded by this library.
"""

from __future__ import annotations

import struct
import zlib
from pathlib import Path
from typing import Sequence, Tuple

import numpy as np

from . import _native

SIGNATURE = b"\x89PNG\r\n\x1a\n"
COLOUR_TYPES = {0: "gray", 2: "RGB", 3: "palette", 4: "gray+alpha", 6: "RGBA"}
_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
_DEPTHS = {0: (1, 2, 4, 8, 16), 2: (8, 16), 3: (1, 2, 4, 8), 4: (8, 16), 6: (8, 16)}
# Adam7 passes: (first row, first column, row step, column step)
_ADAM7 = ((0, 0, 8, 8), (0, 4, 8, 8), (4, 0, 8, 4), (0, 2, 4, 4), (2, 0, 4, 2), (0, 1, 2, 2), (1, 0, 2, 1))


def _chunks(buf: bytes, name: str):
    off = len(SIGNATURE)
    while off + 12 <= len(buf):
        (n,) = struct.unpack_from(">I", buf, off)
        kind = buf[off + 4:off + 8]
        if off + 12 + n > len(buf):
            raise ValueError(f"{name}: chunk {kind!r} runs past the end of the file")
        body = buf[off + 8:off + 8 + n]
        (crc,) = struct.unpack_from(">I", buf, off + 8 + n)
        if zlib.crc32(kind + body) & 0xFFFFFFFF != crc and kind in (b"IHDR", b"PLTE", b"IDAT", b"tRNS"):
            raise ValueError(f"{name}: chunk {kind.decode('latin-1')} fails its CRC; the file is damaged")
        yield kind, body
        off += 12 + n
        if kind == b"IEND":
            return
    raise ValueError(f"{name}: no IEND chunk; the file is truncated")


def png_header(path) -> Tuple[int, int, int, int, int]:
    """(width, height, bit depth, colour type, interlace) from the IHDR,
    refusing a file that is not a PNG or a colour type outside the
    specification (generic_readers.py:54-78)."""
    name = Path(str(path)).name
    with open(str(path), "rb") as f:
        head = f.read(33)
    if len(head) < 26 or head[:8] != SIGNATURE or head[12:16] != b"IHDR":
        raise ValueError(f"{name} does not start with a PNG signature and IHDR chunk, so it is not a PNG file.")
    w, h, depth, ctype, comp, filt, interlace = struct.unpack(">IIBBBBB", head[16:29])
    if ctype not in COLOUR_TYPES:
        raise ValueError(f"{name} has PNG colour type {ctype}, which the PNG specification does not define.")
    if depth not in _DEPTHS[ctype] or comp != 0 or filt != 0 or interlace > 1 or w == 0 or h == 0:
        raise ValueError(f"{name}: IHDR {w}x{h}, depth {depth}, colour type {ctype}, compression {comp}, "
                         f"filter {filt}, interlace {interlace} is not a valid PNG header")
    return w, h, depth, ctype, interlace


def _unfilter(raw: memoryview, rows: int, rowbytes: int, bpp: int, name: str) -> np.ndarray:
    L = _native.core()
    src = np.frombuffer(raw, np.uint8, count=rows * (rowbytes + 1))
    out = np.empty(rows * rowbytes, np.uint8)
    if L.pivamd_png_unfilter(src.ctypes.data, rows, rowbytes, bpp, out.ctypes.data) != 0:
        raise ValueError(f"{name}: a scanline has a filter type above 4; the file is damaged")
    return out.reshape(rows, rowbytes)


def _unpack(rows: np.ndarray, width: int, depth: int, channels: int) -> np.ndarray:
    """Reconstructed scanline bytes -> (rows, width, channels) samples."""
    if depth == 16:
        return rows.view(">u2").reshape(rows.shape[0], -1)[:, :width * channels].astype(np.uint16).reshape(
            rows.shape[0], width, channels)
    if depth == 8:
        return rows[:, :width * channels].reshape(rows.shape[0], width, channels)
    bits = np.unpackbits(rows, axis=1)                       # MSB first, as PNG packs
    bits = bits[:, :width * depth].reshape(rows.shape[0], width, depth)
    weights = (1 << np.arange(depth - 1, -1, -1)).astype(np.uint8)
    return (bits * weights).sum(axis=2, dtype=np.uint8)[:, :, None]


def png_samples(path) -> np.ndarray:
    """The file's samples as stored: (H, W) gray, or (H, W, 3 | 4) colour
    with a palette expanded to its colours (and tRNS alpha)."""
    path = str(path)
    name = Path(path).name
    w, h, depth, ctype, interlace = png_header(path)
    if ctype == 4:
        raise ValueError(f"{name} is a gray+alpha PNG (colour type 4). PIVTOOLs reads gray or colour images and "
                         "does not guess what the alpha channel means for a particle image. Save it as gray, "
                         "without alpha.")
    if ctype == 0 and depth < 8:
        raise ValueError(f"{name} is a {depth}-bit gray PNG; this reader decodes gray PNGs of 8 or 16 bits")
    buf = Path(path).read_bytes()
    idat, plte, trns = [], None, None
    for kind, body in _chunks(buf, name):
        if kind == b"IDAT":
            idat.append(body)
        elif kind == b"PLTE":
            plte = body
        elif kind == b"tRNS":
            trns = body
    if not idat:
        raise ValueError(f"{name}: no IDAT chunk")
    try:
        data = zlib.decompress(b"".join(idat))
    except zlib.error as e:
        raise ValueError(f"{name}: image data damaged ({e})") from None
    ch = _CHANNELS[ctype]
    bpp = max(1, ch * depth // 8)
    out = np.empty((h, w, ch), np.uint16 if depth == 16 else np.uint8)
    passes = _ADAM7 if interlace else ((0, 0, 1, 1),)
    off = 0
    mv = memoryview(data)
    for r0, c0, rs, cs in passes:
        pw, ph = -(-(w - c0) // cs) if w > c0 else 0, -(-(h - r0) // rs) if h > r0 else 0
        if pw == 0 or ph == 0:
            continue
        rowbytes = -(-(pw * ch * depth) // 8)
        need = ph * (rowbytes + 1)
        if off + need > len(data):
            raise ValueError(f"{name}: image data ends early ({len(data)} bytes inflated)")
        rows = _unfilter(mv[off:off + need], ph, rowbytes, bpp, name)
        off += need
        out[r0::rs, c0::cs] = _unpack(rows, pw, depth, ch)
    if ctype == 3:
        if plte is None or len(plte) % 3:
            raise ValueError(f"{name}: palette PNG without a valid PLTE chunk")
        pal = np.frombuffer(plte, np.uint8).reshape(-1, 3)
        idx = out[:, :, 0]
        if int(idx.max()) >= pal.shape[0]:
            raise ValueError(f"{name}: palette index {int(idx.max())} beyond its {pal.shape[0]} colours")
        rgb = pal[idx]
        if trns is not None:
            alpha = np.full(pal.shape[0], 255, np.uint8)
            alpha[:min(len(trns), pal.shape[0])] = np.frombuffer(trns, np.uint8)[:pal.shape[0]]
            rgb = np.concatenate([rgb, alpha[idx][:, :, None]], axis=2)
        return rgb
    return out[:, :, 0] if ch == 1 else out


def read_png(path, device=None):
    """One PNG as a float32 (H, W) device tensor."""
    return read_png_frames([path], device=device)[0]


def read_png_frames(paths: Sequence, device=None, stream_handle=None):
    """Same-shaped PNGs as one float32 (n, H, W) device tensor, one
    conversion launch for all of them."""
    import torch

    paths = list(paths)
    if not paths:
        raise ValueError("no PNG files given")
    dev = torch.device(device) if device is not None else torch.device("cuda")
    frames = [png_samples(p) for p in paths]
    a0 = frames[0]
    for p, a in zip(paths, frames):
        if a.shape != a0.shape or a.dtype != a0.dtype:
            raise ValueError(f"{Path(str(p)).name}: samples {a.shape} {a.dtype} differ from the first file's "
                             f"{a0.shape} {a0.dtype}")
    H, W = a0.shape[:2]
    S = 1 if a0.ndim == 2 else a0.shape[2]
    host = np.ascontiguousarray(np.stack(frames))
    src = torch.from_numpy(host.view(np.uint8).reshape(-1)).to(dev)
    out = torch.empty((len(paths), H, W), dtype=torch.float32, device=dev)
    L = _native.core()
    st = stream_handle if stream_handle is not None else torch.cuda.current_stream(dev).cuda_stream
    code = 1 if a0.dtype == np.uint16 else 0
    _native.check(L.pivamd_gray_from_samples(src.data_ptr(), len(paths), H, W, S, 0, code, out.data_ptr(), st),
                  "gray_from_samples")
    return out

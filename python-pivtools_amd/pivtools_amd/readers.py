"""LaVision DaVis camera containers: ``.im7`` files and ``.set`` recordings.

The reference reads both on the host and hands float frames to the
correlator (``pivtools_core/image_handling/readers/im7_reader.py``,
``set_reader.py``).  Here the host parses the container (headers, manifest,
index, the serial decompressors) and the device does the per-pixel work:
one kernel, ``pivamd_davis_decode``, unpacks a batch of payloads in any of
the seven DaVis pixel encodings, applies the stream's corrections in
LaVision's order and its geometry, and writes the 16-bit frames the pass
loop reads.  The serial codes (the DaVis token stream of ``.im7`` pack 1
and of IMX ``.set`` entries, and LZ4 blocks) run in the library's host code
(``csrc/readers.cpp``); zlib is Python's.

Layouts follow the reference, which settled them against LaVision's lvpyio:

``.set``  (``set_reader.py``)
    ``StreamSet.xml`` declares the frame streams (``ReaderInfo``
    ``Core.Set.Recording.FrameReader`` named ``Images_16bit_8Bit``, bound to
    frame positions ``StartFrame..EndFrame``, contiguous from 0) and the
    transformers bound to each (``:640-1010``).  A stream's
    ``{prefix}-0.ims`` index is a 1024-byte header of int32 words (word 2
    frames per image, 3 width, 4 height; words 0, 1, 5, 14, 15 must hold
    1, 0, 16, 0, 0; bytes 24..55 the codec name) then one ``<Iqq`` entry
    (file number 1, offset, size) per stored frame (``:313-415``);
    ``{prefix}-decoder.xml`` names the encoding (``:422-446``).  Image ``b``
    of a stream spanning ``span`` frame positions is entry
    ``b * span + (position - StartFrame)``.
``.im7``  (``im7_reader.py``)
    A 256-byte ``Image_Header_7`` (``<hhhh iiii hhh``: version, pack type,
    buffer format -3 float / -4 word, sparse flag, sizes x, y, z, f, ...),
    then the pixels: pack 0 raw, 1 DaVis tokens after a 3-byte preamble
    with the running value carried across frames, 2 zlib blocks of one row
    or one frame behind int32 sizes, 20 LZ4 (version 0: one int64-sized
    block over the buffer; version 1: one per frame) (``:110-720``).

Every layout or value the reference refuses is refused here with a
``ValueError`` or ``IOError`` naming it.
"""

from __future__ import annotations

import ctypes
import struct
import xml.etree.ElementTree as ET
import zlib
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _native

# DaVis pixel encodings the device decodes, and their pivamd_davis_decode code
# (set_reader.py:1471-1479).
ENCODINGS = {"mono-10p": 0, "mono-10pmsb": 1, "mono-12p": 2, "mono-12pmsb": 3, "mono-12packed": 4,
             "raw-16-bit": 5, "raw-8-bit": 6}
# Known DaVis encodings that are not decoded (no layout proven), refused by name.
REFUSED_ENCODINGS = ("rgb-24",)
# File-backed corrections in LaVision's application order, with their file
# suffix (set_reader.py:548-558), and the file-less geometries with their
# pivamd_davis_decode code (set_reader.py:560-585; "horinzontal" is DaVis's
# spelling).
CORRECTIONS = {"intensity-correction-1": "scmos-1", "dark-image-subtraction": "dark",
               "background-subtraction": "bgnd"}
GEOMETRIES = {"rotate-180": 1, "vertical-mirror": 2, "horinzontal-mirror": 3, "rotate-cw-90": 4,
              "rotate-ccw-90": 5, "rotate-cw-90-vertical-mirror": 6, "rotate-ccw-90-vertical-mirror": 7}
# Known DaVis ids with no measured effect: refused when their stream is read.
REFUSED_TRANSFORMERS = ("white-image-correction", "add-offset", "image-intensity-correction-2")

FRAME_READER_TYPE = "Core.Set.Recording.FrameReader"
FRAME_READER_NAME = "Images_16bit_8Bit"
IMS_HEADER_BYTES = 1024
_IMS_ENTRY = struct.Struct("<Iqq")
# header word: (field, the only accepted value)
_IMS_REQUIRED_WORDS = {0: ("version", 1), 1: ("image count", 0), 5: ("bits per decoded pixel", 16),
                       14: ("planes per frame", 0), 15: ("sample type", 0)}
IMS_CODECS = ("", "Uncompressed", "Zlib", "LZ4", "IMX")

IM7_HEADER = struct.Struct("<hhhhiiiihhh")
IM7_HEADER_BYTES = 256
BUFFER_FLOAT, BUFFER_WORD = -3, -4
PACK_RAW, PACK_TOKENS, PACK_ZLIB, PACK_FIXED12, PACK_LZ4 = 0, 1, 2, 3, 20
PACK1_PREAMBLE = 3
_INT64 = struct.Struct("<q")
_TOKEN_ERRORS = {1: "the token stream ended inside a token", 2: "a 0x7f run of absolutes runs past the values requested",
                 3: "the running value left the uint16 range 0..65535"}


# ---------------------------------------------------------------------------
# serial decoders (library host code)
# ---------------------------------------------------------------------------

def _buf(data) -> Tuple[np.ndarray, int]:
    a = np.ascontiguousarray(np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data,
                             dtype=np.uint8)
    return a, a.ctypes.data


def davis_tokens(data, n_values: int, offset: int = 0, prev: int = 0) -> Tuple[np.ndarray, int, int]:
    """Decode ``n_values`` DaVis token values from ``data[offset:]`` with the
    running value starting at ``prev`` (davis_tokens.py:46-130).  Returns
    (uint16 values, offset after the last token, last value); raises
    ``IOError`` naming the failure otherwise."""
    a, ptr = _buf(data)
    out = np.empty(n_values, np.uint16)
    off = ctypes.c_size_t(offset)
    pv = ctypes.c_longlong(prev)
    written = ctypes.c_size_t(0)
    st = _native.core().pivamd_davis_tokens(ptr, a.size, ctypes.byref(off), out.ctypes.data, n_values,
                                            ctypes.byref(pv), ctypes.byref(written))
    if st:
        raise IOError(f"{_TOKEN_ERRORS.get(st, f'token status {st}')} (byte {off.value}, value {written.value} "
                      f"of {n_values})")
    return out, off.value, pv.value


def lz4_block(data, expected: int) -> bytes:
    """One LZ4 block decompressed to exactly ``expected`` bytes
    (im7_reader.py:270-291)."""
    a, ptr = _buf(data)
    out = np.empty(max(expected, 1), np.uint8)
    n = _native.core().pivamd_lz4_block(ptr, a.size, out.ctypes.data, expected)
    if n < 0:
        raise IOError("corrupt LZ4 block")
    if n != expected:
        raise IOError(f"LZ4 block decompressed to {n} bytes, expected {expected}")
    return out[:expected].tobytes()


def _zlib_exact(stored: bytes, expected: int, where: str) -> bytes:
    """One bare zlib stream that must inflate to exactly ``expected`` bytes
    with nothing after its end (set_reader.py:1601-1633)."""
    z = zlib.decompressobj()
    try:
        out = z.decompress(stored, expected + 1)
    except zlib.error as exc:
        raise IOError(f"{where}: the zlib stream does not decompress ({exc})") from exc
    if len(out) > expected:
        raise IOError(f"{where}: the zlib stream decompresses to more than {expected} bytes")
    if not z.eof:
        raise IOError(f"{where}: the zlib stream ends after {len(out)} of {expected} bytes without its end marker")
    if len(out) != expected:
        raise IOError(f"{where}: the zlib stream decompresses to {len(out)} bytes, not {expected}")
    if z.unused_data:
        raise IOError(f"{where}: {len(z.unused_data)} bytes follow the end of the zlib stream")
    return out


# ---------------------------------------------------------------------------
# .im7
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class IM7Header:
    version: int
    pack_type: int
    buffer_format: int
    is_sparse: int
    size_x: int
    size_y: int
    size_z: int
    size_f: int


def parse_im7_header(raw: bytes, source: str = ".im7") -> IM7Header:
    """The 256-byte Image_Header_7 (im7_reader.py:149-213)."""
    if len(raw) < IM7_HEADER_BYTES:
        raise IOError(f"{source}: file too small for the {IM7_HEADER_BYTES}-byte header ({len(raw)} bytes)")
    v = IM7_HEADER.unpack_from(raw, 0)
    h = IM7Header(*v[:8])
    bad = {k: getattr(h, k) for k in ("size_x", "size_y", "size_z", "size_f") if getattr(h, k) < 1}
    if bad:
        raise ValueError(f"{source}: damaged .im7 header, every dimension must be at least 1: {bad}")
    if h.version not in (0, 1):
        raise ValueError(f"{source}: .im7 header version {h.version} is not read (0 or 1)")
    if h.version == 1 and h.pack_type != PACK_LZ4:
        raise ValueError(f"{source}: .im7 header version 1 with pack_type {h.pack_type} is not read "
                         f"(version 1 is known with pack_type {PACK_LZ4} only)")
    return h


def _im7_dtype(h: IM7Header, source: str) -> np.dtype:
    if h.buffer_format == BUFFER_FLOAT:
        return np.dtype("<f4")
    if h.buffer_format == BUFFER_WORD:
        return np.dtype("<u2")
    raise ValueError(f"{source}: unsupported buffer_format {h.buffer_format} (-3 float, -4 word)")


def read_im7(path) -> Tuple[IM7Header, np.ndarray]:
    """Header and pixels (size_f, size_z, size_y, size_x) of an .im7 file, in
    its buffer type (uint16 for WORD, float32 for FLOAT)."""
    path = Path(path)
    data = path.read_bytes()
    src = path.name
    h = parse_im7_header(data[:IM7_HEADER_BYTES], src)
    dt = _im7_dtype(h, src)
    npix = h.size_x * h.size_y * h.size_z
    frame_bytes = npix * dt.itemsize
    shape = (h.size_f, h.size_z, h.size_y, h.size_x)
    body = memoryview(data)[IM7_HEADER_BYTES:]
    if h.pack_type == PACK_RAW:
        n = h.size_f * frame_bytes
        if len(body) < n:
            raise IOError(f"{src}: expected {n} bytes of pixel data, got {len(body)}; the file is truncated")
        return h, np.frombuffer(body[:n], dt).reshape(shape).copy()
    if h.pack_type == PACK_TOKENS:
        if h.buffer_format != BUFFER_WORD:
            raise ValueError(f"{src}: pack_type 1 with buffer_format {h.buffer_format} is not read "
                             f"(the token stream codes 16-bit values)")
        arr = np.frombuffer(body, np.uint8)
        out = np.empty((h.size_f, npix), np.uint16)
        off, prev = PACK1_PREAMBLE, 0
        for f in range(h.size_f):
            try:
                out[f], off, prev = davis_tokens(arr, npix, off, prev)
            except IOError as exc:
                raise IOError(f"{src}: pack-1 frame {f} of {h.size_f}: {exc}") from exc
        return h, out.reshape(shape)
    if h.pack_type == PACK_ZLIB:
        row_bytes = h.size_x * dt.itemsize
        pos = 0

        def block_size(p):
            if p + 4 > len(body):
                raise IOError(f"{src}: unexpected end of file reading a zlib block size")
            n = struct.unpack_from("<i", body, p)[0]
            if n <= 0:
                raise IOError(f"{src}: invalid zlib block size {n}")
            return n

        first = block_size(0)
        head = zlib.decompressobj().decompress(bytes(body[4:4 + first]), row_bytes + 1)
        if len(head) > row_bytes:
            block_bytes = frame_bytes
        elif len(head) == row_bytes:
            block_bytes = row_bytes
        else:
            raise IOError(f"{src}: the first zlib block inflates to {len(head)} bytes, not a row ({row_bytes}) "
                          f"or a frame ({frame_bytes})")
        parts = []
        for _ in range(h.size_f * (frame_bytes // block_bytes)):
            n = block_size(pos)
            parts.append(_zlib_exact(bytes(body[pos + 4:pos + 4 + n]), block_bytes, src))
            pos += 4 + n
        return h, np.frombuffer(b"".join(parts), dt).reshape(shape).copy()
    if h.pack_type == PACK_LZ4:
        if h.version == 0:
            n = _INT64.unpack_from(body, 0)[0]
            if n <= 0 or 8 + n > len(body):
                raise IOError(f"{src}: invalid LZ4 compressed size {n}")
            raw = lz4_block(bytes(body[8:8 + n]), h.size_f * frame_bytes)
            return h, np.frombuffer(raw, dt).reshape(shape).copy()
        out = np.empty(shape, dt)
        pos = 0
        starts = []
        for f in range(h.size_f):
            starts.append(IM7_HEADER_BYTES + pos)
            if pos + 8 > len(body):
                raise IOError(f"{src}: version-1 LZ4 frame {f} of {h.size_f} is cut off")
            n = _INT64.unpack_from(body, pos)[0]
            if n <= 0 or pos + 8 + n > len(body):
                raise IOError(f"{src}: version-1 LZ4 frame {f} of {h.size_f} declares {n} compressed bytes")
            out[f] = np.frombuffer(lz4_block(bytes(body[pos + 8:pos + 8 + n]), frame_bytes), dt).reshape(shape[1:])
            pos += 8 + n
        table = np.frombuffer(body[pos:pos + 8 * h.size_f], "<i8")
        if table.size != h.size_f or list(table) != starts:
            raise IOError(f"{src}: the version-1 frame-offset table does not match the frames found")
        return h, out
    if h.pack_type == PACK_FIXED12:
        raise ValueError(f"{src}: pack_type 3 (fixed 12-bit) is not read: no real file exists to check a decoder")
    raise ValueError(f"{src}: pack_type {h.pack_type} is not read (0, 1, 2 or 20)")


def read_im7_frames(path) -> np.ndarray:
    """An .im7 camera file as (size_f, H, W) uint16 counts (one plane per
    frame), refusing float buffers and multi-plane frames."""
    h, px = read_im7(path)
    if h.size_z != 1:
        raise ValueError(f"{Path(path).name}: {h.size_z} planes per frame; a camera image has 1")
    if px.dtype != np.uint16:
        raise ValueError(f"{Path(path).name}: buffer_format {h.buffer_format} is not 16-bit counts")
    return px[:, 0]


def _correction_image(path: Path, stored: Tuple[int, int]) -> np.ndarray:
    """A correction .im7 as uint16 counts in sensor shape (set_reader.py:1062-1138)."""
    _h, px = read_im7(path)
    img = np.squeeze(px)
    if img.ndim != 2:
        raise ValueError(f"correction image {path.name} is {img.ndim}-dimensional; expected one 2D frame")
    lo, hi = img.min(), img.max()
    integral = np.issubdtype(img.dtype, np.integer) or bool(np.all(img == np.floor(img)))
    if not integral or lo < 0 or hi > 65535:
        raise ValueError(f"correction image {path.name} is not integer counts in [0, 65535]")
    if img.shape != tuple(stored):
        raise ValueError(f"correction image {path.name} is {img.shape} but the stream stores {tuple(stored)}")
    return img.astype(np.uint16)


# ---------------------------------------------------------------------------
# .set
# ---------------------------------------------------------------------------

@dataclass
class SetStream:
    """One frame stream of a recording: one frame position of every image."""

    index: int
    prefix: str
    data_path: Path
    decoder: str
    codec: str
    stored_shape: Tuple[int, int]          # (H, W) in sensor coordinates
    frames_per_image: int
    span_offset: int
    entries: List[Tuple[int, int]]         # (offset, size) in the data file
    transformers: Tuple[str, ...] = ()     # ids in application order
    correction_paths: Dict[str, Path] = field(default_factory=dict)

    @property
    def n_images(self) -> int:
        return len(self.entries) // self.frames_per_image

    @property
    def geometry(self) -> int:
        g = [GEOMETRIES[t] for t in self.transformers if t in GEOMETRIES]
        return g[0] if g else 0

    @property
    def shape(self) -> Tuple[int, int]:
        """(H, W) after the geometry: 90-degree turns read as (W, H)."""
        h, w = self.stored_shape
        return (w, h) if self.geometry >= 4 else (h, w)

    def entry(self, image: int) -> Tuple[int, int]:
        if not 0 <= image < self.n_images:
            raise IndexError(f"image {image} out of range [0, {self.n_images})")
        return self.entries[image * self.frames_per_image + self.span_offset]


def _frame_span(node: ET.Element, what: str) -> Tuple[int, int]:
    purpose = node.find("ContentPurpose")
    if purpose is None:
        raise ValueError(f"{what} has no ContentPurpose, so its frame stream is unknown")
    try:
        start, end = int(purpose.get("StartFrame")), int(purpose.get("EndFrame"))
    except (TypeError, ValueError) as exc:
        raise ValueError(f"{what} has a StartFrame/EndFrame that is not a frame-stream number") from exc
    if start < 0 or end < start:
        raise ValueError(f"{what} spans frames {start}..{end}")
    return start, end


def _parse_index(path: Path, frames_per_image: int) -> Tuple[int, int, str, List[Tuple[int, int]]]:
    data = path.read_bytes()
    where = path.name
    if len(data) < IMS_HEADER_BYTES:
        raise ValueError(f"{where} is {len(data)} bytes, shorter than the {IMS_HEADER_BYTES}-byte index header")
    if (len(data) - IMS_HEADER_BYTES) % _IMS_ENTRY.size:
        raise ValueError(f"{where} holds a partial {_IMS_ENTRY.size}-byte entry")
    word = lambda k: struct.unpack_from("<i", data, 4 * k)[0]  # noqa: E731
    for k, (name, value) in _IMS_REQUIRED_WORDS.items():
        if word(k) != value:
            raise ValueError(f"{where}: index header word {k} ({name}) is {word(k)}; only {value} is read")
    if word(2) != frames_per_image:
        raise ValueError(f"{where}: index header word 2 (frames per image) is {word(2)}, but the FrameReader spans "
                         f"{frames_per_image}")
    codec = data[24:56].split(b"\0", 1)[0].decode("latin-1")
    if codec not in IMS_CODECS:
        raise ValueError(f"{where}: compression {codec!r} is not one of {IMS_CODECS}")
    width, height = word(3), word(4)
    entries = []
    for pos in range(IMS_HEADER_BYTES, len(data), _IMS_ENTRY.size):
        file_no, off, size = _IMS_ENTRY.unpack_from(data, pos)
        if file_no != 1:
            raise ValueError(f"{where}: an entry names data file {file_no}; only streams in one data file (1) are read")
        entries.append((off, size))
    if len(entries) % frames_per_image:
        raise ValueError(f"{where}: {len(entries)} entries are not whole images of {frames_per_image} frames")
    return width, height, codec, entries


def _read_decoder(path: Path) -> str:
    if not path.exists():
        raise FileNotFoundError(f"pixel encoding unknown: {path.name} is missing")
    node = ET.parse(path).getroot().find("id")
    if node is None or not (node.text or "").strip():
        raise ValueError(f"pixel encoding unknown: {path} has no <id>")
    return node.text.strip()


class SetRecording:
    """A DaVis ``.set`` recording: its frame streams, read on the device.

    ``path`` is the ``.set`` file or its companion directory (the ``.set``
    file's name without the suffix)."""

    def __init__(self, path):
        p = Path(path)
        self.set_dir = p.with_suffix("") if p.suffix == ".set" else p
        if not self.set_dir.is_dir():
            raise FileNotFoundError(f"companion directory {self.set_dir} not found")
        manifest = self.set_dir / "StreamSet.xml"
        if not manifest.exists():
            raise FileNotFoundError(f"{self.set_dir} has no StreamSet.xml: streams are never guessed from file names")
        root = ET.parse(manifest).getroot()
        owners: Dict[int, Tuple[str, int, int]] = {}
        for node in root.iter("ReaderInfo"):
            if node.get("Type") != FRAME_READER_TYPE:
                continue
            prefix = (node.get("FilePrefix") or "").strip()
            if not prefix:
                raise ValueError(f"a FrameReader in {manifest} has no FilePrefix")
            if node.get("Name") != FRAME_READER_NAME:
                raise ValueError(f"FrameReader '{prefix}' has Name {node.get('Name')!r}, not {FRAME_READER_NAME!r}")
            start, end = _frame_span(node, f"FrameReader '{prefix}'")
            for f in range(start, end + 1):
                if f in owners:
                    raise ValueError(f"two image streams are bound to frame stream {f}")
                owners[f] = (prefix, start, end)
        if not owners:
            raise ValueError(f"{manifest} declares no FrameReader")
        if sorted(owners) != list(range(len(owners))):
            raise ValueError(f"{manifest} declares frame streams {sorted(owners)}; expected contiguous from 0")
        transformers = self._transformers(root, manifest)
        self.streams: List[SetStream] = []
        for i in range(len(owners)):
            prefix, start, end = owners[i]
            span = end - start + 1
            width, height, codec, entries = _parse_index(self.set_dir / f"{prefix}-0.ims", span)
            decoder = _read_decoder(self.set_dir / f"{prefix}-decoder.xml")
            ids, paths = transformers.get(i, ((), {}))
            self.streams.append(SetStream(i, prefix, self.set_dir / f"{prefix}-1.ims", decoder, codec,
                                          (height, width), span, i - start, entries, ids, paths))
        counts = {s.n_images for s in self.streams}
        if len(counts) != 1:
            raise ValueError(f"frame streams hold different image counts: {sorted(counts)}")
        self.n_images = counts.pop()
        self._corr_cache: Dict[Tuple[int, str], object] = {}

    def _transformers(self, root, manifest):
        declared: Dict[int, List[Tuple[str, str]]] = {}
        for node in root.iter("Transformer"):
            tid = (node.get("ID") or "").strip()
            prefix = (node.get("FilePrefix") or "").strip()
            start, end = _frame_span(node, f"Transformer '{prefix}'")
            if start != end:
                raise ValueError(f"Transformer '{prefix}' spans frames {start}..{end}; one frame is read")
            on = declared.setdefault(start, [])
            for other, _p in on:
                if other == tid:
                    raise ValueError(f"'{tid}' is declared twice for frame stream {start}")
                if tid in GEOMETRIES and other in GEOMETRIES:
                    raise ValueError(f"two geometric transformers for frame stream {start}: '{other}' and '{tid}'")
            on.append((tid, prefix))
        out = {}
        rank = {t: k for k, t in enumerate([*CORRECTIONS, *GEOMETRIES, *REFUSED_TRANSFORMERS])}
        for stream, items in declared.items():
            for tid, _p in items:
                if tid not in rank:
                    raise ValueError(f"frame stream {stream} declares the unknown transformer '{tid}'")
            items = sorted(items, key=lambda t: rank[t[0]])
            paths = {}
            for tid, prefix in items:
                if tid in CORRECTIONS:
                    path = self.set_dir / f"{prefix}-{CORRECTIONS[tid]}.im7"
                    if not path.exists():
                        raise FileNotFoundError(f"'{tid}' of frame stream {stream} needs {path.name}, which is missing")
                    paths[tid] = path
            out[stream] = (tuple(t for t, _ in items), paths)
        return out

    # -- payloads ---------------------------------------------------------------

    def payload(self, stream: int, image: int) -> bytes:
        """The packed payload of one image of one stream (decompressed)."""
        s = self.streams[stream]
        if s.decoder in REFUSED_ENCODINGS or s.decoder not in ENCODINGS:
            raise ValueError(f"pixel encoding '{s.decoder}' is not decoded; decoded: {', '.join(ENCODINGS)}")
        h, w = s.stored_shape
        expected = int(_native.core().pivamd_davis_payload_bytes(ENCODINGS[s.decoder], h, w))
        if expected == 0:
            raise ValueError(f"a {w}x{h} frame does not fill whole groups of encoding '{s.decoder}'")
        off, size = s.entry(image)
        with open(s.data_path, "rb") as f:
            f.seek(off)
            stored = f.read(size)
        if len(stored) != size:
            raise IOError(f"{s.data_path.name}: expected {size} bytes at offset {off}, got {len(stored)}")
        where = f"frame stream {stream}, image {image} ({s.codec or 'uncompressed'})"
        if s.codec in ("", "Uncompressed"):
            raw = stored
        elif s.codec == "Zlib":
            raw = _zlib_exact(stored, expected, where)
        elif s.codec == "LZ4":
            raw = lz4_block(stored, expected)
        else:   # IMX: tokens over the payload's little-endian 16-bit words
            if expected % 2:
                raise ValueError(f"{where}: IMX codes 16-bit words, but the payload is {expected} bytes")
            words, end, _ = davis_tokens(stored, expected // 2)
            if end != len(stored):
                raise IOError(f"{where}: IMX tokens end at byte {end}, the entry holds {len(stored)}")
            raw = words.astype("<u2").tobytes()
        if len(raw) != expected:
            raise ValueError(f"{where}: {len(raw)} payload bytes, encoding '{s.decoder}' needs {expected}")
        return raw

    def read_frames(self, stream: int, images, device=None, stream_handle=None):
        """Images of one stream as an int16 device tensor (n, H, W) holding
        the unsigned counts, decoded, corrected and turned on the device."""
        import torch
        s = self.streams[stream]
        for t in s.transformers:
            if t in REFUSED_TRANSFORMERS:
                raise ValueError(f"frame stream {stream} declares '{t}', whose effect is unmeasured; not applied")
        images = list(images)
        dev = torch.device(device) if device is not None else torch.device("cuda")
        h, w = s.stored_shape
        payloads = [self.payload(stream, b) for b in images]
        nb = len(payloads[0]) if payloads else 0
        host = torch.empty((len(images), max(nb, 1)), dtype=torch.uint8, pin_memory=True)
        for k, p in enumerate(payloads):
            host[k, :nb] = torch.frombuffer(bytearray(p), dtype=torch.uint8)
        raw = host.to(dev, non_blocking=True)
        corr = [self._correction(stream, t, dev) for t in CORRECTIONS]
        ho, wo = s.shape
        out = torch.empty((len(images), ho, wo), dtype=torch.int16, device=dev)
        st = stream_handle if stream_handle is not None else torch.cuda.current_stream(dev).cuda_stream
        ptr = lambda t: 0 if t is None else t.data_ptr()  # noqa: E731
        if images:
            _native.check(_native.core().pivamd_davis_decode(
                raw.data_ptr(), nb, len(images), h, w, ENCODINGS[s.decoder], ptr(corr[0]), ptr(corr[1]),
                ptr(corr[2]), s.geometry, out.data_ptr(), st), "davis_decode")
        return out

    def _correction(self, stream: int, tid: str, dev):
        import torch
        s = self.streams[stream]
        if tid not in s.transformers:
            return None
        key = (stream, tid)
        if key not in self._corr_cache:
            img = _correction_image(s.correction_paths[tid], s.stored_shape)
            self._corr_cache[key] = torch.from_numpy(img.view(np.int16).copy()).to(dev)
        return self._corr_cache[key]

"""CPU oracle of camera-payload decoding -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of the reference's frame decoding:
  unpack_mono12p        pivtools_core/image_handling/readers/set_reader.py:1296-1350
                        (PFNC Mono12p, two 12-bit pixels in three bytes,
                        LSB first), in the byte-plane form
  subtract_dark         pivraw_reader.py:937-955 (max(p, d) - d per photosite
                        in the sensor type)
  bayer_rg8_superpixel  pivraw_reader.py:956-959 (dark per photosite, then the
                        2x2 superpixel sum)
  dark_role             pivraw_reader.py:916-934 (a double-frame dark holds
                        A then B; 1-based odd frames are A)

Pins (tests/test_ingest_oracle.py):
  - the reference's `.pivraw` contract fixtures: payloads written by the
    acquisition app's recorder and the frames that app decodes from them,
    clean and dark-subtracted, Mono12p and BayerRG8
    (tests/golden/pivraw_vectors.npz);
  - the reference test's hand-decoded bytes (0x34 0x12 0xAB -> 564, 2737);
  - a round trip through an independent packer that writes the LSB-first bit
    stream bit by bit (numpy.packbits, bitorder='little');
  - the dark clamp values of the reference's dark test.
"""

from __future__ import annotations

import numpy as np


def unpack_mono12p(raw, width: int, height: int) -> np.ndarray:
    """Mono12p payload -> uint16 (height, width)."""
    n = int(width) * int(height)
    if n % 2:
        raise ValueError(f"Mono12p needs an even pixel count, got {width}x{height}")
    b = np.frombuffer(bytes(raw), dtype=np.uint8) if not isinstance(raw, np.ndarray) else raw.reshape(-1)
    if b.size < n // 2 * 3:
        raise ValueError(f"{b.size} bytes for {n} Mono12p pixels")
    b = b[: n // 2 * 3].astype(np.uint16)
    b0, b1, b2 = b[0::3], b[1::3], b[2::3]
    out = np.empty(n, dtype=np.uint16)
    out[0::2] = b0 | ((b1 & 0x0F) << 8)
    out[1::2] = (b1 >> 4) | (b2 << 4)
    return out.reshape(height, width)


def pack_mono12p_bits(pixels: np.ndarray) -> np.ndarray:
    """Independent packer for the round trip: the LSB-first bit stream of
    12-bit values, written bit by bit."""
    v = np.asarray(pixels, dtype=np.uint16).reshape(-1)
    if v.size % 2:
        raise ValueError("Mono12p packs pixels in pairs")
    bits = ((v[:, None] >> np.arange(12, dtype=np.uint16)[None, :]) & 1).astype(np.uint8).reshape(-1)
    return np.packbits(bits, bitorder="little")


def subtract_dark(pixels: np.ndarray, dark: np.ndarray | None) -> np.ndarray:
    """max(p, d) - d in the sensor's unsigned type (never wraps)."""
    if dark is None:
        return pixels.copy()
    p = pixels.copy()
    np.maximum(p, dark.astype(p.dtype), out=p)
    p -= dark.astype(p.dtype)
    return p


def bayer_rg8_superpixel(mosaic: np.ndarray, dark: np.ndarray | None = None) -> np.ndarray:
    """(H, W) uint8 mosaic -> (H/2, W/2) sums of the 2x2 superpixels of the
    dark-subtracted photosites (exact integers; the reference holds them as
    float32)."""
    H, W = mosaic.shape
    if H % 2 or W % 2:
        raise ValueError("BayerRG8 superpixels need an even mosaic")
    p = subtract_dark(mosaic.astype(np.uint8), None if dark is None else dark.astype(np.uint8))
    return p.reshape(H // 2, 2, W // 2, 2).sum(axis=(1, 3), dtype=np.uint16)


def dark_role(frame_1based: int, dark_mode: str) -> int:
    """Index of the dark frame for a 1-based file frame."""
    return (frame_1based - 1) % 2 if dark_mode == "double_frame" else 0


# ---------------------------------------------------------------------------
# DaVis .set / .im7 containers
# ---------------------------------------------------------------------------
# Restated from pivtools_core/image_handling/readers/set_reader.py:1196-1480
# (pixel encodings), :477-585 (corrections and geometries, LaVision's order),
# davis_tokens.py:46-130 (the token stream) and im7_reader.py:270-430 (LZ4).
# Pins (tests/test_ingest_oracle.py): LaVision's lvpyio frames for every
# geometry, every correction and their orders (the reference's
# lvpyio_oracle.npz recordings), for mono-10pmsb / mono-12pmsb /
# mono-12packed payloads, for .im7 pack 1 and pack 20 files and an IMX .set
# stream; the reference tests' hand-computed bytes; round trips through
# independent bit-stream packers; the LZ4 block format against zlib-free
# hand-built blocks.

def _bytes(raw) -> np.ndarray:
    return np.frombuffer(bytes(raw), dtype=np.uint8) if not isinstance(raw, np.ndarray) else raw.reshape(-1)


def decode_davis(raw, encoding: str, width: int, height: int) -> np.ndarray:
    """A DaVis payload -> uint16 (height, width), by byte planes."""
    n = int(width) * int(height)
    b = _bytes(raw).astype(np.uint32)
    out = np.empty(n, dtype=np.uint32)
    if encoding in ("mono-10p", "mono-10pmsb"):
        if n % 4:
            raise ValueError(f"{encoding} packs four pixels in five bytes; {width}x{height}")
        g = b[: n // 4 * 5].reshape(-1, 5)
        if encoding == "mono-10p":       # little-endian bit stream
            out[0::4] = g[:, 0] | (g[:, 1] & 0x03) << 8
            out[1::4] = g[:, 1] >> 2 | (g[:, 2] & 0x0F) << 6
            out[2::4] = g[:, 2] >> 4 | (g[:, 3] & 0x3F) << 4
            out[3::4] = g[:, 3] >> 6 | g[:, 4] << 2
        else:                            # big-endian bit stream
            out[0::4] = g[:, 0] << 2 | g[:, 1] >> 6
            out[1::4] = (g[:, 1] & 0x3F) << 4 | g[:, 2] >> 4
            out[2::4] = (g[:, 2] & 0x0F) << 6 | g[:, 3] >> 2
            out[3::4] = (g[:, 3] & 0x03) << 8 | g[:, 4]
    elif encoding in ("mono-12p", "mono-12pmsb", "mono-12packed"):
        if n % 2:
            raise ValueError(f"{encoding} packs two pixels in three bytes; {width}x{height}")
        g = b[: n // 2 * 3].reshape(-1, 3)
        if encoding == "mono-12p":
            out[0::2] = g[:, 0] | (g[:, 1] & 0x0F) << 8
            out[1::2] = g[:, 1] >> 4 | g[:, 2] << 4
        elif encoding == "mono-12pmsb":
            out[0::2] = g[:, 0] << 4 | g[:, 1] >> 4
            out[1::2] = (g[:, 1] & 0x0F) << 8 | g[:, 2]
        else:                            # nibble byte in the middle
            out[0::2] = g[:, 0] << 4 | (g[:, 1] & 0x0F)
            out[1::2] = g[:, 2] << 4 | g[:, 1] >> 4
    elif encoding == "raw-16-bit":
        out[:] = b[0:2 * n:2] | b[1:2 * n:2] << 8
    elif encoding == "raw-8-bit":
        out[:] = b[:n]
    else:
        raise ValueError(f"encoding {encoding!r} is not decoded")
    return out.astype(np.uint16).reshape(height, width)


def pack_bits(pixels, bits: int, msb_first: bool) -> np.ndarray:
    """Independent packer: the pixels' bits as one continuous stream, LSB-first
    (PFNC 'p') or MSB-first ('pmsb'), written bit by bit."""
    v = np.asarray(pixels, dtype=np.uint32).reshape(-1)
    k = np.arange(bits, dtype=np.uint32)
    order = k[::-1] if msb_first else k
    stream = ((v[:, None] >> order[None, :]) & 1).astype(np.uint8).reshape(-1)
    return np.packbits(stream, bitorder="big" if msb_first else "little")


def correct_davis(img: np.ndarray, scmos=None, dark=None, bgnd=None) -> np.ndarray:
    """LaVision's corrections in their order, each clipped to its range:
    clip(x - S + 300, 0, 65535), then max(x - D, 0), then max(x - B, 0)."""
    x = img.astype(np.int64)
    if scmos is not None:
        x = np.clip(x - scmos.astype(np.int64) + 300, 0, 65535)
    if dark is not None:
        x = np.maximum(x - dark.astype(np.int64), 0)
    if bgnd is not None:
        x = np.maximum(x - bgnd.astype(np.int64), 0)
    return x.astype(np.uint16)


def geometry_davis(img: np.ndarray, geometry) -> np.ndarray:
    """The output frame of one geometry, by explicit index maps
    (out[i, j] = x[r(i, j), c(i, j)])."""
    H, W = img.shape
    if geometry is None:
        return img.copy()
    turns = geometry.startswith("rotate-cw-90") or geometry.startswith("rotate-ccw-90")
    Ho, Wo = (W, H) if turns else (H, W)
    i, j = np.meshgrid(np.arange(Ho), np.arange(Wo), indexing="ij")
    r, c = {
        "rotate-180": (H - 1 - i, W - 1 - j),
        "vertical-mirror": (H - 1 - i, j),
        "horinzontal-mirror": (i, W - 1 - j),
        "rotate-cw-90": (H - 1 - j, i),
        "rotate-ccw-90": (j, W - 1 - i),
        "rotate-cw-90-vertical-mirror": (H - 1 - j, W - 1 - i),
        "rotate-ccw-90-vertical-mirror": (j, i),
    }[geometry]
    return img[r, c]


def davis_tokens(data, n_values: int, offset: int = 0, prev: int = 0):
    """The DaVis token grammar, one token at a time (pure Python: small inputs).
    Returns (uint16 values, offset after the last token, last value)."""
    d = bytes(data)
    out = []
    off = offset
    while len(out) < n_values:
        if off >= len(d):
            raise IOError("truncated")
        b = d[off]
        off += 1
        if b == 0x81:
            while len(out) < n_values:
                if off >= len(d):
                    raise IOError("truncated")
                byte = d[off]
                off += 1
                hi, lo = byte >> 4, byte & 0xF
                prev += hi - 16 if hi >= 8 else hi
                if not 0 <= prev <= 65535:
                    raise IOError("range")
                out.append(prev)
                if len(out) >= n_values or lo == 8:
                    break
                prev += lo - 16 if lo >= 8 else lo
                if not 0 <= prev <= 65535:
                    raise IOError("range")
                out.append(prev)
        elif b == 0x80:
            if off + 2 > len(d):
                raise IOError("truncated")
            prev = d[off] | d[off + 1] << 8
            off += 2
            out.append(prev)
        elif b == 0x7F:
            if off >= len(d):
                raise IOError("truncated")
            cnt = d[off]
            off += 1
            if len(out) + cnt > n_values:
                raise IOError("overrun")
            if off + 2 * cnt > len(d):
                raise IOError("truncated")
            for _ in range(cnt):
                prev = d[off] | d[off + 1] << 8
                off += 2
                out.append(prev)
        else:
            prev += b - 256 if b >= 128 else b
            if not 0 <= prev <= 65535:
                raise IOError("range")
            out.append(prev)
    return np.array(out, dtype=np.uint16), off, prev


def lz4_block(data, expected: int) -> bytes:
    """The LZ4 block format, one sequence at a time (pure Python)."""
    d = bytes(data)
    out = bytearray()
    ip = 0
    while ip < len(d):
        token = d[ip]
        ip += 1
        lit = token >> 4
        if lit == 15:
            while True:
                s = d[ip]
                ip += 1
                lit += s
                if s != 255:
                    break
        out += d[ip:ip + lit]
        ip += lit
        if ip >= len(d):
            break
        off = d[ip] | d[ip + 1] << 8
        ip += 2
        if off == 0 or off > len(out):
            raise IOError("bad match offset")
        ml = (token & 15) + 4
        if token & 15 == 15:
            while True:
                s = d[ip]
                ip += 1
                ml += s
                if s != 255:
                    break
        for _ in range(ml):
            out.append(out[-off])
    if len(out) != expected:
        raise IOError(f"{len(out)} bytes, expected {expected}")
    return bytes(out)

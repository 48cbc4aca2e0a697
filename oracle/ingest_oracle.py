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

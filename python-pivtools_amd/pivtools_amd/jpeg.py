"""JPEG frames: decoded on the host, made gray on the device.

Follows the JPEG half of the reference's ``read_png_jpeg``
(``pivtools_core/image_handling/readers/generic_readers.py:125-176``): OpenCV's
``imread`` with ``IMREAD_UNCHANGED`` decodes with libjpeg-turbo at its
defaults and applies no EXIF orientation; a colour file becomes BT.601 luma
by OpenCV's RGB2GRAY rule, a gray file keeps its samples, and the frame is
float32.

The decoder is the library's own (``csrc/jpeg.cpp``, host code: the entropy
decoding is serial), with libjpeg-turbo's default arithmetic so that its
samples are the same bytes; the luma and the float conversion run on the
device (``pivamd_gray_from_samples``, shared with the TIFF and PNG readers).
Progressive, lossless and arithmetic-coded files, 12-bit samples, CMYK and
unusual chroma sampling are refused by name.
"""

from __future__ import annotations

from pathlib import Path
from typing import Sequence

import numpy as np

from . import _native

_STATUS = {1: "is damaged or truncated", 2: "is a progressive JPEG, which this reader does not decode",
           3: "uses arithmetic coding, which this reader does not decode",
           4: "is a lossless or hierarchical JPEG, which this reader does not decode",
           5: "has 12-bit samples; this reader decodes 8-bit JPEG",
           6: "has 2 or 4 colour components (CMYK / YCCK); this reader decodes gray and colour JPEG",
           7: "uses chroma sampling other than 4:4:4, 4:2:2 or 4:2:0", 8: "needs a larger output buffer"}


def jpeg_samples(path) -> np.ndarray:
    """(H, W) gray or (H, W, 3) RGB uint8 samples of a JPEG file."""
    path = str(path)
    name = Path(path).name
    data = np.frombuffer(Path(path).read_bytes(), np.uint8)
    L = _native.core()
    info = np.zeros(3, np.int32)
    rc = int(L.pivamd_jpeg_info(data.ctypes.data, data.size, info.ctypes.data))
    if rc:
        raise ValueError(f"{name} {_STATUS.get(rc, f'fails to decode ({rc})')}")
    w, h, c = (int(v) for v in info)
    out = np.empty((h, w, c), np.uint8)
    rc = int(L.pivamd_jpeg_decode(data.ctypes.data, data.size, out.ctypes.data, out.size))
    if rc:
        raise ValueError(f"{name} {_STATUS.get(rc, f'fails to decode ({rc})')}")
    return out[:, :, 0] if c == 1 else out


def read_jpeg(path, device=None):
    """One JPEG as a float32 (H, W) device tensor."""
    return read_jpeg_frames([path], device=device)[0]


def read_jpeg_frames(paths: Sequence, device=None, stream_handle=None):
    """Same-shaped JPEGs as one float32 (n, H, W) device tensor, one
    conversion launch for all of them."""
    import torch

    paths = list(paths)
    if not paths:
        raise ValueError("no JPEG files given")
    dev = torch.device(device) if device is not None else torch.device("cuda")
    frames = [jpeg_samples(p) for p in paths]
    a0 = frames[0]
    for p, a in zip(paths, frames):
        if a.shape != a0.shape:
            raise ValueError(f"{Path(str(p)).name}: samples {a.shape} differ from the first file's {a0.shape}")
    H, W = a0.shape[:2]
    S = 1 if a0.ndim == 2 else a0.shape[2]
    host = np.ascontiguousarray(np.stack(frames))
    src = torch.from_numpy(host.reshape(-1)).to(dev)
    out = torch.empty((len(paths), H, W), dtype=torch.float32, device=dev)
    L = _native.core()
    st = stream_handle if stream_handle is not None else torch.cuda.current_stream(dev).cuda_stream
    _native.check(L.pivamd_gray_from_samples(src.data_ptr(), len(paths), H, W, S, 0, 0, out.data_ptr(), st),
                  "gray_from_samples")
    return out

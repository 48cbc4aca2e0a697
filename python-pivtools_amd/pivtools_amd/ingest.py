"""Camera payloads decoded on the device.

The reference decodes each raw frame on the host before correlating it
(pivtools_core/image_handling/readers/set_reader.py:1296-1350 decode_mono12p;
pivraw_reader.py:896-961 _sensor_pixels / _decode_into).  Here the packed
payload is what is copied to the GPU -- 1.5 bytes per pixel for Mono12p
instead of 2 for 16-bit frames -- and one kernel expands a whole batch in
HBM into the 16-bit frames that ``MultipassPIV`` and ``EnsemblePIV`` take.

Formats (the ones the reference's `.pivraw` container records):

* ``mono12p``: PFNC Mono12p, two 12-bit pixels in three bytes, LSB first;
* ``bayer_rg8``: an 8-bit RG Bayer mosaic read as 2x2 superpixels (sum of the
  four photosites), so an H x W sensor gives an H/2 x W/2 frame.

A dark frame, in the sensor's own type, is subtracted per photosite as
``max(p, d) - d`` before anything else, as the reference does.  Outputs are
int16 tensors holding the unsigned counts (at most 4095 and 1020), the
16-bit layout the correlation kernels read.

Reading the container files themselves (headers, index, drop alignment) is
host I/O and stays with the caller: these functions take the payload bytes.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native

FORMATS = ("mono12p", "bayer_rg8")


def payload_bytes(fmt: str, height: int, width: int) -> int:
    """Bytes of one frame's payload for a sensor of height x width."""
    if fmt == "mono12p":
        if (height * width) % 2:
            raise ValueError(f"Mono12p packs two pixels into three bytes; {width}x{height} is an odd pixel count")
        return height * width * 3 // 2
    if fmt == "bayer_rg8":
        if height % 2 or width % 2:
            raise ValueError(f"BayerRG8 superpixels need an even mosaic, got {width}x{height}")
        return height * width
    raise ValueError(f"unknown payload format {fmt!r}; expected one of {FORMATS}")


def frame_shape(fmt: str, height: int, width: int) -> tuple:
    """Shape of the decoded frame for a sensor of height x width."""
    payload_bytes(fmt, height, width)
    return (height // 2, width // 2) if fmt == "bayer_rg8" else (height, width)


def decode(fmt: str, raw: torch.Tensor, height: int, width: int, *, n_frames: int | None = None,
           frame_stride: int | None = None, dark: torch.Tensor | None = None,
           out: torch.Tensor | None = None) -> torch.Tensor:
    """Decode n_frames payloads of a height x width sensor.

    ``raw`` is a uint8 device tensor; frame f's payload starts at byte
    ``f * frame_stride`` (default: the payloads packed back to back, or the
    row length of a 2-D ``raw``).  ``dark`` is a device tensor of one dark
    frame in the sensor type (uint16/int16 H x W for Mono12p, uint8 H x W for
    BayerRG8) or None.  Returns (n_frames, h, w) int16 (``out`` if given).
    """
    nbytes = payload_bytes(fmt, height, width)
    if raw.dtype != torch.uint8 or not raw.is_cuda:
        raise TypeError("payloads are uint8 device tensors")
    if not raw.is_contiguous():
        raise ValueError("payload tensor must be contiguous")
    if frame_stride is None:
        frame_stride = raw.shape[1] if raw.dim() == 2 else nbytes
    if n_frames is None:
        n_frames = raw.shape[0] if raw.dim() == 2 else max(1, (raw.numel() - nbytes) // frame_stride + 1)
    n_frames = int(n_frames)
    if frame_stride < nbytes and n_frames > 1:
        raise ValueError(f"frame stride {frame_stride} is shorter than one payload ({nbytes} bytes)")
    need = (n_frames - 1) * frame_stride + nbytes if n_frames else 0
    if raw.numel() < need:
        raise ValueError(f"{raw.numel()} payload bytes for {n_frames} frames need {need}")
    h, w = frame_shape(fmt, height, width)
    if out is None:
        out = torch.empty((n_frames, h, w), dtype=torch.int16, device=raw.device)
    elif (tuple(out.shape) != (n_frames, h, w) or out.dtype not in (torch.int16, torch.uint16)
          or not out.is_contiguous() or out.device != raw.device):
        raise ValueError(f"out must be a contiguous 16-bit ({n_frames}, {h}, {w}) tensor on {raw.device}")
    if dark is not None:
        want = (torch.int16, torch.uint16) if fmt == "mono12p" else (torch.uint8,)
        if dark.dtype not in want or dark.numel() != height * width or not dark.is_contiguous() \
                or dark.device != raw.device:
            raise ValueError(f"dark must be one contiguous {height}x{width} sensor-type frame on {raw.device}")
    if n_frames == 0:
        return out
    L = _native.core()
    st = torch.cuda.current_stream(raw.device).cuda_stream
    fn = L.pivamd_unpack_mono12p if fmt == "mono12p" else L.pivamd_bayer_rg8_superpixel
    # frames of one call go in blocks of at most 65535 (the launch's y extent)
    for f0 in range(0, n_frames, 65535):
        nf = min(65535, n_frames - f0)
        rc = fn(raw.data_ptr() + f0 * frame_stride, frame_stride, nf, height, width,
                None if dark is None else dark.data_ptr(), out[f0:f0 + nf].data_ptr(), st)
        _native.check(rc, f"{fmt} decode")
    return out


def unpack_mono12p(raw, height, width, **kw) -> torch.Tensor:
    return decode("mono12p", raw, height, width, **kw)


def bayer_rg8_superpixel(raw, height, width, **kw) -> torch.Tensor:
    return decode("bayer_rg8", raw, height, width, **kw)


class PayloadFeeder:
    """Pinned host staging for payload batches.

    ``put(frames)`` copies a (n, nbytes) uint8 host array into pinned memory
    and from there to the device on the current stream without blocking the
    host; ``decode`` then expands it into 16-bit frames on the same stream.
    Two staging sets alternate, so the host can fill the next batch while the
    previous copy is in flight.
    """

    def __init__(self, fmt: str, height: int, width: int, max_frames: int, device="cuda"):
        self.fmt, self.height, self.width = fmt, int(height), int(width)
        self.nbytes = payload_bytes(fmt, self.height, self.width)
        self.dev = torch.device(device)
        self.max_frames = int(max_frames)
        self._host = [torch.empty((self.max_frames, self.nbytes), dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self._dev = [torch.empty((self.max_frames, self.nbytes), dtype=torch.uint8, device=self.dev) for _ in range(2)]
        self._done = [None, None]
        self._k = 0

    def put(self, frames: np.ndarray) -> torch.Tensor:
        frames = np.ascontiguousarray(frames, dtype=np.uint8).reshape(-1, self.nbytes)
        n = frames.shape[0]
        if n > self.max_frames:
            raise ValueError(f"{n} frames exceed the feeder's {self.max_frames}")
        k = self._k
        self._k ^= 1
        if self._done[k] is not None:
            self._done[k].synchronize()       # the copy that last used this pinned set has finished
        self._host[k][:n].numpy()[...] = frames
        d = self._dev[k][:n]
        d.copy_(self._host[k][:n], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.dev))
        self._done[k] = ev
        return d

    def decode(self, payloads: torch.Tensor, dark: torch.Tensor | None = None,
               out: torch.Tensor | None = None) -> torch.Tensor:
        return decode(self.fmt, payloads, self.height, self.width, dark=dark, out=out)

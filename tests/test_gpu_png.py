"""PNG frames made gray on the device, against the reference's own PNG tests
(unit-tests/test_generic_readers.py:136-222): gray16 keeps every value,
colour16 with equal channels reads as the gray16 file, (40000, 20000, 60000)
within 1.6 of 30540, (100, 50, 200) -> 82 for RGB and palette files, and
on the seed-0 512 x 512 noise image exactly 301 pixels differ from Pillow's
convert("L"), all by 1, each within 0.51 of exact BT.601."""

import os
import sys

import numpy as np
import pytest
import torch
from PIL import Image

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_png import PNG_GRAY, PNG_PALETTE, PNG_RGB, PNG_RGBA, write_png  # noqa: E402

from pivtools_amd import png as P  # noqa: E402

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")
H, W = 16, 24


def _read(path):
    out = P.read_png(path, device=dev)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _ramp16():
    return np.linspace(0, 65535, H * W).reshape(H, W).astype(np.uint16)


def test_gray16_png_keeps_every_value(tmp_path):
    out = _read(write_png(tmp_path / "gray16.png", _ramp16(), PNG_GRAY, filters=False))
    assert out.dtype == np.float32
    np.testing.assert_array_equal(out, _ramp16().astype(np.float32))


@pytest.mark.parametrize("ctype", [PNG_RGB, PNG_RGBA], ids=["rgb16", "rgba16"])
def test_colour16_with_equal_channels_reads_as_the_gray16_file(tmp_path, ctype):
    ramp = _ramp16()
    ch = [ramp, ramp, ramp] + ([np.full_like(ramp, 65535)] if ctype == PNG_RGBA else [])
    out = _read(write_png(tmp_path / "c16.png", np.dstack(ch), ctype, filters=False))
    assert out.max() == 65535
    np.testing.assert_array_equal(out, _read(write_png(tmp_path / "g16.png", ramp, PNG_GRAY, filters=False)))


def test_rgb16_reads_as_full_depth_luma(tmp_path):
    rgb = np.empty((H, W, 3), np.uint16)
    rgb[...] = (40000, 20000, 60000)
    np.testing.assert_allclose(_read(write_png(tmp_path / "rgb16.png", rgb, PNG_RGB)), 30540.0, atol=1.6)


def test_rgb8_and_palette_read_as_the_tiff_luma(tmp_path):
    rgb = np.empty((H, W, 3), np.uint8)
    rgb[...] = (100, 50, 200)
    np.testing.assert_array_equal(_read(write_png(tmp_path / "rgb8.png", rgb, PNG_RGB)),
                                  np.full((H, W), 82.0, np.float32))
    idx = np.zeros((H, W), np.uint8)
    np.testing.assert_array_equal(_read(write_png(tmp_path / "pal.png", idx, PNG_PALETTE, palette=[(100, 50, 200)])),
                                  np.full((H, W), 82.0, np.float32))


def test_rgb8_luma_is_bt601_and_differs_from_pillow_by_rounding_only(tmp_path):
    rgb = np.random.default_rng(0).integers(0, 256, (512, 512, 3), np.uint8)
    path = write_png(tmp_path / "noise.png", rgb, PNG_RGB, filters=False)
    exact = rgb.astype(np.float64) @ np.array([0.299, 0.587, 0.114])
    pillow = np.array(Image.open(path).convert("L")).astype(np.float32)
    out = _read(path)
    assert np.abs(out - exact).max() <= 0.51
    assert np.abs(out - pillow).max() == 1
    assert int((out != pillow).sum()) == 301


def test_batch_of_pngs(tmp_path):
    paths = [write_png(tmp_path / f"p{k}.png", np.random.default_rng(k).integers(0, 256, (30, 40, 3), np.uint8),
                       PNG_RGB, interlace=bool(k % 2)) for k in range(3)]
    out = P.read_png_frames(paths, device=dev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), np.stack([_read(p) for p in paths]))

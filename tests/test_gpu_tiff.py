"""TIFF frames made gray on the device (pivamd_gray_from_samples), against
what the reference's own tests state of its read_tiff / _luma
(unit-tests/test_generic_readers.py:26-35, 75-84, 161-222):

* (100, 50, 200) reads as 82 for contiguous and planar pages;
* 8-bit colour: within 0.51 of exact BT.601, and against Pillow's
  convert("L") on the seed-0 512 x 512 noise image exactly 301 pixels
  differ, all by 1 -- the count the reference measured with OpenCV 4.12,
  which pins OpenCV's fixed-point rule;
* 16-bit colour within 1.6 counts of exact BT.601 (measured there: 1.52
  over 262144 random pixels), (40000, 20000, 60000) -> 30540;
* gray pages keep every value; other sample counts average."""

import os
import sys

import numpy as np
import pytest
import torch
from PIL import Image

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_tiff import write_tiff  # noqa: E402

from pivtools_amd import tiff as T  # noqa: E402

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")
BT601 = np.array([0.299, 0.587, 0.114])


def _read(path):
    out = T.read_tiff(path, device=dev)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("planar", [False, True])
def test_uniform_rgb_reads_as_its_luma(tmp_path, planar):
    rgb = np.empty((16, 24, 3), np.uint8)
    rgb[...] = (100, 50, 200)
    data = np.moveaxis(rgb, -1, 0).copy() if planar else rgb
    out = _read(write_tiff(tmp_path / "rgb.tif", data, planar=planar))
    assert out.dtype == np.float32 and out.shape == (16, 24)
    np.testing.assert_array_equal(out, np.full((16, 24), 82.0, np.float32))


def test_rgb8_luma_is_opencv_rule_against_pillow(tmp_path):
    rgb = np.random.default_rng(0).integers(0, 256, (512, 512, 3), np.uint8)
    path = str(tmp_path / "noise.tif")
    Image.fromarray(rgb).save(path)
    out = _read(path)
    exact = rgb.astype(np.float64) @ BT601
    pillow = np.array(Image.fromarray(rgb).convert("L")).astype(np.float32)
    assert np.abs(out - exact).max() <= 0.51
    assert np.abs(out - pillow).max() == 1
    assert int((out != pillow).sum()) == 301


def test_rgb16_luma_within_fixed_point_error(tmp_path):
    rgb = np.empty((16, 24, 3), np.uint16)
    rgb[...] = (40000, 20000, 60000)
    out = _read(write_tiff(tmp_path / "u.tif", rgb))
    np.testing.assert_allclose(out, 30540.0, atol=1.6)
    rnd = np.random.default_rng(1).integers(0, 65536, (512, 512, 3)).astype(np.uint16)
    out = _read(write_tiff(tmp_path / "r.tif", rnd, compression=8, rows_per_strip=64))
    assert np.abs(out - rnd.astype(np.float64) @ BT601).max() <= 1.6


def test_rgba_ignores_alpha_and_two_samples_average(tmp_path):
    rgb = np.random.default_rng(2).integers(0, 256, (20, 30, 3), np.uint8)
    rgba = np.concatenate([rgb, np.random.default_rng(3).integers(0, 256, (20, 30, 1), np.uint8)], axis=2)
    np.testing.assert_array_equal(_read(write_tiff(tmp_path / "a.tif", rgba, extra_tags=[(338, 3, [2])])),
                                  _read(write_tiff(tmp_path / "b.tif", rgb)))
    ga = np.random.default_rng(4).integers(0, 65536, (20, 30, 2)).astype(np.uint16)
    out = _read(write_tiff(tmp_path / "g.tif", ga, extra_tags=[(338, 3, [2])]))
    np.testing.assert_array_equal(out, np.mean(ga, axis=-1).astype(np.float32))
    gf = np.random.default_rng(5).normal(size=(20, 30, 5)).astype(np.float32)
    out = _read(write_tiff(tmp_path / "f.tif", gf))
    np.testing.assert_array_equal(out, np.mean(gf, axis=-1).astype(np.float32))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.int32, np.float32])
def test_gray_pages_keep_every_value(tmp_path, dtype):
    rng = np.random.default_rng(6)
    info = np.iinfo(dtype) if np.dtype(dtype).kind in "ui" else None
    data = (rng.integers(info.min, info.max, (37, 41), endpoint=True) if info else rng.normal(size=(37, 41)))
    data = data.astype(dtype)
    out = _read(write_tiff(tmp_path / "g.tif", data, tile=(16, 16), compression=8, predictor=2 if info else 1))
    np.testing.assert_array_equal(out, data.astype(np.float32))


def test_float_rgb_luma_in_float(tmp_path):
    rgb = np.random.default_rng(7).uniform(0, 1000, (20, 30, 3)).astype(np.float32)
    out = _read(write_tiff(tmp_path / "f.tif", rgb))
    r, g, b = (rgb[..., k] for k in range(3))
    want = (r * np.float32(0.299) + g * np.float32(0.587)) + b * np.float32(0.114)
    np.testing.assert_array_equal(out, want)


def test_batch_of_frames_in_one_launch(tmp_path):
    paths, want = [], []
    for k in range(3):
        rgb = np.random.default_rng(10 + k).integers(0, 256, (40, 56, 3), np.uint8)
        paths.append(write_tiff(tmp_path / f"f{k}.tif", rgb, compression=8))
        want.append(_read(paths[-1]))
    out = T.read_tiff_frames(paths, device=dev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), np.stack(want))
    other = write_tiff(tmp_path / "o.tif", np.zeros((40, 56), np.uint8))
    with pytest.raises(ValueError, match=r"o\.tif.*differ"):
        T.read_tiff_frames(paths + [other], device=dev)

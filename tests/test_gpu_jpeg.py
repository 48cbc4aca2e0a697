"""JPEG frames made gray on the device, against the reference's own JPEG
tests (unit-tests/test_generic_readers.py:228-258): the samples are Pillow's
decode (OpenCV's at those versions), colour through BT.601 within 0.51 of the
exact luma, gray exactly; an EXIF orientation is not applied."""

import numpy as np
import pytest
import torch
from PIL import Image

from pivtools_amd import jpeg as J

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _random_rgb8(seed=0, side=64):
    return np.random.default_rng(seed).integers(0, 256, (side, side, 3), np.uint8)


@pytest.mark.parametrize("colour", [True, False], ids=["colour", "gray"])
def test_jpeg_reads_as_its_decoded_samples(tmp_path, colour):
    data = _random_rgb8() if colour else _random_rgb8()[..., 1]
    path = str(tmp_path / "frame.jpg")
    Image.fromarray(data).save(path, quality=90)
    decoded = np.array(Image.open(path)).astype(np.float64)
    if colour:
        decoded = decoded @ np.array([0.299, 0.587, 0.114])
    out = J.read_jpeg(path, device=dev)
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    assert out.shape == (64, 64) and out.dtype == np.float32
    assert np.abs(out - decoded).max() <= (0.51 if colour else 0)


def test_jpeg_exif_orientation_is_not_applied(tmp_path):
    exif = Image.Exif()
    exif[0x0112] = 6
    path = str(tmp_path / "rotated.jpg")
    Image.fromarray(np.zeros((16, 24), np.uint8)).save(path, exif=exif.tobytes())
    assert J.read_jpeg(path, device=dev).shape == (16, 24)


def test_batch_of_jpegs(tmp_path):
    paths = []
    for k in range(3):
        p = str(tmp_path / f"f{k}.jpg")
        Image.fromarray(_random_rgb8(k, 48)).save(p, quality=85, subsampling=k)
        paths.append(p)
    out = J.read_jpeg_frames(paths, device=dev)
    torch.cuda.synchronize()
    for k, p in enumerate(paths):
        rgb = np.array(Image.open(p)).astype(np.int64)
        want = (rgb[..., 0] * 9798 + rgb[..., 1] * 19235 + rgb[..., 2] * 3735 + 16384) >> 15
        np.testing.assert_array_equal(out[k].cpu().numpy(), want.astype(np.float32))

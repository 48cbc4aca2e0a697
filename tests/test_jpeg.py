"""JPEG decoding on the host against Pillow's libjpeg-turbo decode, which the
reference's tests take as the truth for OpenCV's
(unit-tests/test_generic_readers.py:228-258): the same bytes for gray and
colour files at every chroma sampling Pillow writes, odd and tiny sizes,
optimised Huffman tables and restart intervals.  Refusals name what they
refuse.  The gray conversion runs on the device (tests/test_gpu_jpeg.py)."""

import numpy as np
import pytest
from PIL import Image

from pivtools_amd import _native
from pivtools_amd import jpeg as J


def _roundtrip(tmp_path, img, mode, **kw):
    path = str(tmp_path / "f.jpg")
    Image.fromarray(img, mode).save(path, **kw)
    return J.jpeg_samples(path), np.array(Image.open(path))


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (9, 17), (16, 16), (31, 7), (53, 77), (120, 96)])
@pytest.mark.parametrize("sub", [0, 1, 2])
def test_colour_equals_pillow(tmp_path, shape, sub):
    rng = np.random.default_rng(shape[0] * 100 + shape[1] + sub)
    img = rng.integers(0, 256, shape + (3,), np.uint8)
    if shape[0] >= 2:                                    # a flat half for the zero-AC shortcuts
        img[: shape[0] // 2] = img[: shape[0] // 2].mean(axis=(0, 1), keepdims=True).astype(np.uint8)
    ours, pil = _roundtrip(tmp_path, img, "RGB", quality=85, subsampling=sub)
    assert ours.shape == shape + (3,)
    np.testing.assert_array_equal(ours, pil)


@pytest.mark.parametrize("quality", [30, 75, 95, 100])
def test_gray_equals_pillow(tmp_path, quality):
    rng = np.random.default_rng(quality)
    x = np.linspace(0, 200, 67)[None, :] + np.linspace(0, 50, 45)[:, None] + rng.normal(0, 12, (45, 67))
    ours, pil = _roundtrip(tmp_path, x.clip(0, 255).astype(np.uint8), "L", quality=quality)
    assert ours.shape == (45, 67) and ours.dtype == np.uint8
    np.testing.assert_array_equal(ours, pil)


@pytest.mark.parametrize("kw", [{"optimize": True}, {"restart_marker_blocks": 1}, {"restart_marker_blocks": 5},
                                {"restart_marker_rows": 2}])
def test_huffman_tables_and_restart_intervals(tmp_path, kw):
    img = np.random.default_rng(9).integers(0, 256, (70, 90, 3), np.uint8)
    ours, pil = _roundtrip(tmp_path, img, "RGB", quality=80, **kw)
    np.testing.assert_array_equal(ours, pil)


def test_refusals_name_what_they_refuse(tmp_path):
    img = np.random.default_rng(1).integers(0, 256, (32, 32, 3), np.uint8)
    Image.fromarray(img).save(tmp_path / "prog.jpg", progressive=True)
    with pytest.raises(ValueError, match=r"prog\.jpg is a progressive JPEG"):
        J.jpeg_samples(tmp_path / "prog.jpg")
    Image.fromarray(img).convert("CMYK").save(tmp_path / "cmyk.jpg")
    with pytest.raises(ValueError, match=r"cmyk\.jpg has 2 or 4 colour components"):
        J.jpeg_samples(tmp_path / "cmyk.jpg")
    raw = (tmp_path / "prog.jpg").read_bytes()
    (tmp_path / "junk.jpg").write_bytes(b"\x00" + raw[1:])
    with pytest.raises(ValueError, match=r"junk\.jpg is damaged"):
        J.jpeg_samples(tmp_path / "junk.jpg")
    Image.fromarray(img).save(tmp_path / "ok.jpg")
    cut = (tmp_path / "ok.jpg").read_bytes()
    (tmp_path / "cut.jpg").write_bytes(cut[:300])
    with pytest.raises(ValueError, match=r"cut\.jpg is damaged"):
        J.jpeg_samples(tmp_path / "cut.jpg")


def test_decoder_bounds_on_hostile_input():
    """Random bytes behind a valid header never read or write out of bounds
    (every status is a refusal or a whole frame)."""
    L = _native.core()
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (24, 40, 3), np.uint8)
    import io
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG", quality=90)
    good = np.frombuffer(buf.getvalue(), np.uint8)
    out = np.zeros(24 * 40 * 3 + 64, np.uint8)
    for k in range(200):
        bad = good.copy()
        idx = rng.integers(2, bad.size, 8)
        bad[idx] = rng.integers(0, 256, 8)
        info = np.zeros(3, np.int32)
        rc = L.pivamd_jpeg_info(bad.ctypes.data, bad.size, info.ctypes.data)
        if rc == 0 and int(info[0]) * int(info[1]) * int(info[2]) <= 24 * 40 * 3:
            L.pivamd_jpeg_decode(bad.ctypes.data, bad.size, out.ctypes.data, 24 * 40 * 3)
    assert not out[24 * 40 * 3:].any()

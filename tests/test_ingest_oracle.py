"""CPU checks of the camera-payload oracle (no GPU).

The vectors come from the reference's `.pivraw` contract fixtures: payloads
written by the acquisition app's recorder and the frames that app decodes,
clean and dark-subtracted (tests/golden/pivraw_vectors.npz, made by
tests/golden/make_pivraw_vectors.py).
"""

import json
import os

import numpy as np
import pytest

import ingest_oracle as IO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pivraw_vectors.npz")


def _vectors():
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta_json"]).decode())
    return z, meta


def decode_case(z, meta, key, dark: bool):
    """Decode every frame of a fixture case with the oracle."""
    m = meta[key]
    frames = z[key + "__payload"]
    dpay = z[key + "__dark_payload"] if dark else None
    out = []
    for k in range(frames.shape[0]):
        if m["pixel_format"] == "Mono12p":
            p = IO.unpack_mono12p(frames[k], m["width"], m["height"])
            d = None
            if dark:
                d = IO.unpack_mono12p(dpay[IO.dark_role(k + 1, m["dark_mode"])], m["width"], m["height"])
            out.append(IO.subtract_dark(p, d))
        else:
            mosaic = frames[k].reshape(m["height"], m["width"])
            d = dpay[IO.dark_role(k + 1, m["dark_mode"])].reshape(m["height"], m["width"]) if dark else None
            out.append(IO.bayer_rg8_superpixel(mosaic, d))
    return np.stack(out)


@pytest.mark.parametrize("key", ["df_2cam_clean__Camera1", "df_2cam_clean__Camera2", "df_2cam_dark__Camera1",
                                 "df_2cam_dark__Camera2", "tr_bayer_clean__Camera1", "tr_bayer_dark__Camera1"])
def test_oracle_decodes_the_recorder_fixtures(key):
    z, meta = _vectors()
    np.testing.assert_array_equal(decode_case(z, meta, key, False), z[key + "__decoded"])
    if key + "__dark_payload" in z.files:
        got = decode_case(z, meta, key, True)
        np.testing.assert_array_equal(got, z[key + "__dark_subtracted"])
        # the dark actually changes something, and clamps at zero
        assert (got != z[key + "__decoded"]).any()


def test_bayer_dark_is_per_photosite_before_the_sum():
    """A dark subtracted per photosite differs from one subtracted after the
    2x2 sum on the recorder's Bayer fixture (the reference's own argument)."""
    z, meta = _vectors()
    key = "tr_bayer_dark__Camera1"
    m = meta[key]
    mosaic = z[key + "__payload"][0].reshape(m["height"], m["width"])
    dark = z[key + "__dark_payload"][0].reshape(m["height"], m["width"])
    per = IO.bayer_rg8_superpixel(mosaic, dark).astype(np.int64)
    first = np.clip(IO.bayer_rg8_superpixel(mosaic).astype(np.int64) - IO.bayer_rg8_superpixel(dark), 0, None)
    np.testing.assert_array_equal(per, z[key + "__dark_subtracted"][0])
    assert (per != first).any()


def test_hand_decoded_bytes():
    """0x34 0x12 0xAB -> 0x234, 0xAB1 (the reference's hand-decoded pair)."""
    np.testing.assert_array_equal(IO.unpack_mono12p(bytes([0x34, 0x12, 0xAB]), 2, 1), [[564, 2737]])
    np.testing.assert_array_equal(IO.unpack_mono12p(bytes([0xFF] * 3), 2, 1), [[4095, 4095]])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_round_trip_through_an_independent_packer(seed):
    rng = np.random.default_rng(seed)
    h, w = 6, 10
    px = rng.integers(0, 4096, size=(h, w), dtype=np.uint16)
    raw = IO.pack_mono12p_bits(px)
    assert raw.size == h * w * 3 // 2
    np.testing.assert_array_equal(IO.unpack_mono12p(raw, w, h), px)


def test_bit_patterns_against_the_packer():
    """Walking bits: one set bit of the stream lands in exactly one pixel."""
    for bit in range(48):
        raw = np.zeros(6, dtype=np.uint8)
        raw[bit // 8] = 1 << (bit % 8)
        got = IO.unpack_mono12p(raw, 4, 1).reshape(-1)
        want = np.zeros(4, dtype=np.uint16)
        want[bit // 12] = 1 << (bit % 12)
        np.testing.assert_array_equal(got, want)


def test_dark_clamps_at_zero():
    """The reference's dark test values: 10-20 and 0-20 clamp to 0."""
    raw = np.array([[100, 50, 10], [200, 30, 0]], dtype=np.uint16)
    dark = np.full((2, 3), 20, dtype=np.uint16)
    np.testing.assert_array_equal(IO.subtract_dark(raw, dark), [[80, 30, 0], [180, 10, 0]])


def test_odd_pixel_count_is_refused():
    with pytest.raises(ValueError):
        IO.unpack_mono12p(bytes(3), 3, 1)

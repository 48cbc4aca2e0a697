"""DaVis containers on the device: pivamd_davis_decode (decode, corrections
and geometry in one pass) against LaVision's lvpyio frames and the oracle."""

import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import ingest_oracle as I  # noqa: E402
from pivtools_amd import _native  # noqa: E402
from pivtools_amd import readers as R  # noqa: E402
from test_davis_readers import GOLD, _transformer_cases, _unpack  # noqa: E402

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _u16(t):
    return t.cpu().numpy().view(np.uint16)


@pytest.mark.parametrize("case", _transformer_cases())
def test_device_set_read_against_lvpyio(case, tmp_path):
    """Every geometry, correction and order of the reference's lvpyio oracle,
    read through SetRecording.read_frames (one device launch for all images)."""
    z = np.load(os.path.join(GOLD, "transformers_lvpyio.npz"))
    rec = R.SetRecording(_unpack(z, case, str(tmp_path)))
    ref = z[f"{case}__lvpyio"]
    out = rec.read_frames(0, range(rec.n_images))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_u16(out), ref)


def test_device_set_imx_against_lvpyio(tmp_path):
    z = np.load(os.path.join(GOLD, "set_imx_lvpyio.npz"))
    for k in z.files:
        if k == "lvpyio":
            continue
        p = os.path.join(str(tmp_path), k.replace("~", "/"))
        os.makedirs(os.path.dirname(p), exist_ok=True)
        z[k].tofile(p)
    rec = R.SetRecording(os.path.join(str(tmp_path), "fixture.set"))
    out = _u16(rec.read_frames(0, [0]))[0]
    np.testing.assert_array_equal(out, z["lvpyio"][: out.shape[0], : out.shape[1]])


def _decode(raw: np.ndarray, n, H, W, enc, geo, scmos=None, dark=None, bgnd=None, stride=None):
    L = _native.core()
    code = R.ENCODINGS[enc]
    nb = int(L.pivamd_davis_payload_bytes(code, H, W))
    stride = stride or nb
    d_raw = torch.as_tensor(raw, device=dev)
    corr = [None if c is None else torch.as_tensor(c.view(np.int16), device=dev) for c in (scmos, dark, bgnd)]
    turns = geo >= 4
    out = torch.full((n, W, H) if turns else (n, H, W), -1, dtype=torch.int16, device=dev)
    ptr = lambda t: 0 if t is None else t.data_ptr()  # noqa: E731
    rc = L.pivamd_davis_decode(d_raw.data_ptr(), stride, n, H, W, code, ptr(corr[0]), ptr(corr[1]), ptr(corr[2]),
                               geo, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, _u16(out)


@pytest.mark.parametrize("enc", ["mono-10pmsb", "mono-12pmsb", "mono-12packed"])
def test_device_packed_against_lvpyio(enc):
    z = np.load(os.path.join(GOLD, "packed_lvpyio.npz"))
    key = enc.replace("-", "_")
    rc, out = _decode(z[key + "__payload"], 1, 48, 64, enc, 0)
    assert rc == 0
    np.testing.assert_array_equal(out[0], z[key + "__lvpyio"])


@pytest.mark.parametrize("enc", list(R.ENCODINGS))
def test_device_every_encoding_geometry_correction(enc):
    """Each encoding on random payloads of 3 frames at a padded stride, with
    each geometry and all corrections (values that clip both ways), against
    the oracle."""
    rng = np.random.default_rng(R.ENCODINGS[enc])
    H, W, n = 12, 20, 3
    L = _native.core()
    nb = int(L.pivamd_davis_payload_bytes(R.ENCODINGS[enc], H, W))
    stride = nb + 7
    raw = rng.integers(0, 256, n * stride).astype(np.uint8)
    scmos = rng.integers(0, 700, (H, W)).astype(np.uint16)
    dark = rng.integers(0, 300, (H, W)).astype(np.uint16)
    bgnd = rng.integers(0, 300, (H, W)).astype(np.uint16)
    geos = [None, *R.GEOMETRIES]
    for gi, g in enumerate(geos):
        corr = [(None, None, None), (scmos, dark, bgnd)][gi % 2]
        rc, out = _decode(raw, n, H, W, enc, 0 if g is None else R.GEOMETRIES[g], *corr, stride=stride)
        assert rc == 0
        for f in range(n):
            want = I.geometry_davis(I.correct_davis(I.decode_davis(raw[f * stride:f * stride + nb], enc, W, H),
                                                    *corr), g)
            np.testing.assert_array_equal(out[f], want, err_msg=f"{enc} {g} frame {f}")


def test_device_decode_refusals():
    L = _native.core()
    st = torch.cuda.current_stream().cuda_stream
    out = torch.empty(64, dtype=torch.int16, device=dev)
    raw = torch.zeros(64, dtype=torch.uint8, device=dev)
    # mono-10p needs whole groups of four pixels
    assert L.pivamd_davis_decode(raw.data_ptr(), 10, 1, 1, 3, 0, 0, 0, 0, 0, out.data_ptr(), st) == 2
    assert L.pivamd_davis_decode(raw.data_ptr(), 10, 1, 2, 2, 9, 0, 0, 0, 0, out.data_ptr(), st) == 1
    assert L.pivamd_davis_decode(raw.data_ptr(), 10, 1, 2, 2, 5, 0, 0, 0, 8, out.data_ptr(), st) == 1

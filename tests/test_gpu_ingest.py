"""Camera payloads decoded on the device against the oracle and the
recorder's own fixtures.

* the reference's `.pivraw` contract fixtures (Mono12p double-frame with an
  A/B dark, BayerRG8 time-resolved with a dark), placed back at the
  container's 4096-byte frame stride: bit-identical to the frames the
  acquisition app decodes;
* random payloads at full camera size and at ragged sizes, in the fast
  (aligned, whole 8-pixel groups) and the byte-wise forms (odd strides,
  unaligned bases, tails), with and without a dark: bit-identical to the
  oracle;
* size errors come back as error codes;
* a serving stream fed Mono12p payloads gives the same fields as a 16-bit
  stream fed the oracle's decode of the same payloads.
"""

import json
import os

import numpy as np
import pytest
import torch

import ingest_oracle as IO
from pivtools_amd import _native, ingest

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pivraw_vectors.npz")


def _strided(payloads: np.ndarray, stride: int, offset: int = 0) -> tuple:
    """Payload rows placed at `stride` bytes in one device buffer, starting
    `offset` bytes in; returns (buffer view from the first payload, stride)."""
    n, nb = payloads.shape
    buf = np.zeros(offset + max(0, n - 1) * stride + nb + 16, np.uint8)
    for k in range(n):
        buf[offset + k * stride: offset + k * stride + nb] = payloads[k]
    d = torch.from_numpy(buf).to(dev)
    return d[offset:], stride


@pytest.mark.parametrize("key", ["df_2cam_clean__Camera1", "df_2cam_clean__Camera2", "df_2cam_dark__Camera1",
                                 "df_2cam_dark__Camera2", "tr_bayer_clean__Camera1", "tr_bayer_dark__Camera1"])
def test_recorder_fixtures(key):
    z = np.load(GOLDEN, allow_pickle=False)
    m = json.loads(bytes(z["meta_json"]).decode())[key]
    print(f"case {key}", flush=True)
    fmt = "mono12p" if m["pixel_format"] == "Mono12p" else "bayer_rg8"
    pay = z[key + "__payload"]
    raw, stride = _strided(pay, m["frame_stride"])
    got = ingest.decode(fmt, raw, m["height"], m["width"], n_frames=pay.shape[0], frame_stride=stride)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint16), z[key + "__decoded"])
    if key + "__dark_payload" not in z.files:
        return
    dpay = z[key + "__dark_payload"]
    want = z[key + "__dark_subtracted"]
    for k in range(pay.shape[0]):
        role = IO.dark_role(k + 1, m["dark_mode"])
        if fmt == "mono12p":
            dark = torch.from_numpy(IO.unpack_mono12p(dpay[role], m["width"], m["height"]).view(np.int16)).to(dev)
        else:
            dark = torch.from_numpy(dpay[role].copy()).to(dev)
        one, _ = _strided(pay[k:k + 1], m["frame_stride"])
        got = ingest.decode(fmt, one, m["height"], m["width"], n_frames=1, dark=dark)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy()[0].view(np.uint16), want[k])


def _mono_case(n, H, W, stride, offset, with_dark, seed):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 4096, size=(n, H, W), dtype=np.uint16)
    pay = np.stack([IO.pack_mono12p_bits(px[k]) for k in range(n)]) if H * W <= 4096 else None
    if pay is None:
        # large frames: pack with the closed form (pinned to the bit packer on small ones)
        flat = px.reshape(n, -1).astype(np.uint32)
        p0, p1 = flat[:, 0::2], flat[:, 1::2]
        pay = np.empty((n, H * W // 2, 3), np.uint8)
        pay[..., 0] = p0 & 0xFF
        pay[..., 1] = ((p0 >> 8) & 0xF) | ((p1 & 0xF) << 4)
        pay[..., 2] = p1 >> 4
        pay = pay.reshape(n, -1)
    dark = rng.integers(0, 300, size=(H, W), dtype=np.uint16) if with_dark else None
    raw, st = _strided(pay, stride, offset)
    d_dark = torch.from_numpy(dark.view(np.int16)).to(dev) if with_dark else None
    got = ingest.decode("mono12p", raw, H, W, n_frames=n, frame_stride=st, dark=d_dark)
    torch.cuda.synchronize()
    want = np.stack([IO.subtract_dark(IO.unpack_mono12p(pay[k], W, H), dark) for k in range(n)])
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint16), want)
    np.testing.assert_array_equal(want, IO.subtract_dark(px, dark))


@pytest.mark.parametrize("n,H,W,stride,offset", [
    (2, 2160, 2560, 8294400 + 4096, 0),                        # full frames, padded stride (fast form)
    (3, 64, 48, 64 * 48 * 3 // 2, 0),                          # packed back to back, whole groups
    (3, 6, 5, 45, 0),                                          # 30 px: a 6-pixel tail, odd stride
    (2, 16, 16, 390, 1),                                       # unaligned base and stride
    (1, 1, 2, 3, 0),                                           # one pixel pair
])
@pytest.mark.parametrize("with_dark", [False, True])
def test_mono12p_against_oracle(n, H, W, stride, offset, with_dark):
    print(f"mono12p n={n} {H}x{W} stride={stride} offset={offset} dark={with_dark}", flush=True)
    _mono_case(n, H, W, stride, offset, with_dark, seed=H * 7 + W + n)


@pytest.mark.parametrize("n,H,W,stride,offset", [
    (2, 1080, 1280, 1080 * 1280, 0),       # fast form
    (3, 10, 6, 61, 0),                     # W % 8 != 0, odd stride
    (2, 8, 16, 136, 3),                    # unaligned base
])
@pytest.mark.parametrize("with_dark", [False, True])
def test_bayer_against_oracle(n, H, W, stride, offset, with_dark):
    print(f"bayer n={n} {H}x{W} stride={stride} offset={offset} dark={with_dark}", flush=True)
    rng = np.random.default_rng(H + W + n)
    mos = rng.integers(0, 256, size=(n, H, W), dtype=np.uint8)
    dark = rng.integers(0, 40, size=(H, W), dtype=np.uint8) if with_dark else None
    raw, st = _strided(mos.reshape(n, -1), stride, offset)
    got = ingest.decode("bayer_rg8", raw, H, W, n_frames=n, frame_stride=st,
                        dark=torch.from_numpy(dark).to(dev) if with_dark else None)
    torch.cuda.synchronize()
    want = np.stack([IO.bayer_rg8_superpixel(mos[k], dark) for k in range(n)])
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint16), want)


def test_size_errors_are_codes():
    L = _native.core()
    buf = torch.zeros(64, dtype=torch.uint8, device=dev)
    out = torch.zeros(64, dtype=torch.int16, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    assert L.pivamd_unpack_mono12p(buf.data_ptr(), 8, 1, 3, 3, None, out.data_ptr(), st) == 2    # odd count
    assert L.pivamd_unpack_mono12p(buf.data_ptr(), 5, 2, 2, 4, None, out.data_ptr(), st) == 1    # stride < payload
    assert L.pivamd_bayer_rg8_superpixel(buf.data_ptr(), 16, 1, 3, 4, None, out.data_ptr(), st) == 2
    assert L.pivamd_unpack_mono12p(buf.data_ptr(), 8, 0, 2, 4, None, out.data_ptr(), st) == 0    # nothing to do
    with pytest.raises(ValueError):
        ingest.decode("mono12p", buf, 3, 3)


def test_serving_stream_of_mono12p_payloads():
    """A Mono12p stream with A/B darks equals a 16-bit stream fed the
    oracle's decode of the same payloads, pair for pair."""
    from pivtools_amd import synthetic as S
    from pivtools_amd.serving import PIVServer, StreamSpec

    shape = (256, 320)
    rng = np.random.default_rng(3)
    dark_a = rng.integers(0, 40, size=shape, dtype=np.uint16)
    dark_b = rng.integers(0, 40, size=shape, dtype=np.uint16)
    specs = [StreamSpec("u16", shape, [(64, 64), (32, 32)], [50, 50]),
             StreamSpec("m12", shape, [(64, 64), (32, 32)], [50, 50], pixel_format="mono12p", sensor_shape=shape,
                        dark_a=torch.from_numpy(dark_a.view(np.int16)).to(dev),
                        dark_b=torch.from_numpy(dark_b.view(np.int16)).to(dev))]
    srv = PIVServer(specs, max_batch=4, max_slots_per_stream=2)
    reqs = []
    for k in range(6):
        a, b = S.multipass_pair(seed=900 + k, shape=shape, device=dev)
        pa = IO.pack_mono12p_bits(np.clip(a.cpu().numpy().view(np.uint16), 0, 4095))
        pb = IO.pack_mono12p_bits(np.clip(b.cpu().numpy().view(np.uint16), 0, 4095))
        ua = IO.subtract_dark(IO.unpack_mono12p(pa, shape[1], shape[0]), dark_a)
        ub = IO.subtract_dark(IO.unpack_mono12p(pb, shape[1], shape[0]), dark_b)
        r16 = srv.submit(0, torch.from_numpy(ua.view(np.int16)).to(dev), torch.from_numpy(ub.view(np.int16)).to(dev))
        r12 = srv.submit(1, torch.from_numpy(pa).to(dev), torch.from_numpy(pb).to(dev))
        reqs.append((r16, r12))
        srv.poll()
    srv.drain()
    for r16, r12 in reqs:
        x16, x12 = r16.result(), r12.result()
        for u, v in zip(x16, x12):
            assert np.array_equal(u.cpu().numpy(), v.cpu().numpy(), equal_nan=True)

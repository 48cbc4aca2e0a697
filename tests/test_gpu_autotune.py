"""Every launch variant of the correlation kernel against the oracle, and
the autotuner's selection and cache.

Each FP32 FFT variant (R lanes per line transform) must meet the same bar
as the default build: identical NaN masks and integer peaks, planes within
1e-5 relative L2 of the oracle, displacements within 0.005 px.  (The MFMA
variant has its own tests in test_gpu_split_dft.py.)
"""

import json

import numpy as np
import pytest
import torch

import oracle as O
from gpu_util import dropin_bulkxcorr2d
from pivtools_amd import autotune as AT
from pivtools_amd.window import compute_window_centers, window_weight
from test_gpu_dropin import _check_outputs, _pair

pytestmark = pytest.mark.gpu

# the FP32 variants; the split-precision MFMA forms have their own bars in test_gpu_split_dft
CASES = [(n, v) for n in (8, 12, 16, 24, 32, 48, 64, 96, 128) for v in AT.variants(n) if v not in AT.MFMA_FORMS]


@pytest.mark.parametrize("n,v", CASES)
def test_variant_matches_oracle(n, v):
    shape = (3 * n + 7, 4 * n + 5)
    a, b = _pair(n + 1000 * v, shape)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[1, 0] = 1.0
    print(f"variant {v} of {n}x{n}", flush=True)
    prev = AT.set_variant(n, v)
    try:
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    finally:
        AT.set_variant(n, prev)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    _check_outputs(got, ref)
    assert np.all(got["pk_x"][0, 0, 1, 0] == 0.0)     # masked window untouched


def test_tuner_selects_and_caches(tmp_path):
    cache = str(tmp_path / "variants.json")
    sizes = [16, 48]
    try:
        table = AT.tune_xcorr(sizes, reps=3, cache=cache)
        for n in sizes:
            rec = table[n]
            assert rec["variant"] in AT.variants(n)
            assert AT.get_variant(n) == rec["variant"]
            # every variant but bf16 hi/lo (about 16 bits: selectable, never tuned in)
            assert set(rec["ms"]) == {str(v) for v in AT.variants(n) if v != AT.GEMM + 1}
            assert rec["rejected"] == []                  # every variant reproduces the default
            fastest = min(rec["ms"].values())
            chosen = rec["ms"][str(rec["variant"])]
            assert chosen == fastest or (rec["variant"] == rec["default"] and fastest > 0.97 * chosen)
        stored = json.load(open(cache))
        assert len(stored) == 1
        # second call: taken from the cache, no measurement
        for n in sizes:
            AT.set_variant(n, 0)
        again = AT.tune_xcorr(sizes, reps=3, cache=cache)
        assert again == {n: json.loads(json.dumps(table[n])) for n in sizes}
        assert all(AT.get_variant(n) == table[n]["variant"] for n in sizes)
    finally:
        for n in sizes:
            AT.set_variant(n, 0)
    torch.cuda.synchronize()


def test_tune_warp_picks_a_tile_shape_with_equal_frames():
    """The warp tuner times both FMA tile shapes, finds their frames equal,
    and leaves one of them selected."""
    from pivtools_amd import autotune
    rec = autotune.tune_warp((300, 520), n_pairs=2, reps=2)
    assert rec["equal"] is True
    assert rec["variant"] in (0, 1) and set(rec["ms"]) == {"0", "1"}
    assert autotune.get_warp_variant() == rec["variant"]
    autotune.set_warp_variant(1)

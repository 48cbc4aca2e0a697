"""The FP64 plane reference of bench.py's mixed-window accuracy report
against the oracle's FP64 planes (no GPU: torch float64 on the CPU)."""

import os
import sys

import numpy as np
import pytest
import torch

import oracle as O
from pivtools_amd import synthetic as S
from pivtools_amd.window import compute_window_centers, window_weight

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("n,shape", [(16, (100, 133)), (48, (203, 250)), (64, (256, 256))])
def test_planes_fp64_matches_oracle(n, shape):
    rng = np.random.default_rng(n)
    x, y, inten = S._positions(rng, shape, 0.03)
    a = S.render(shape, x, y, inten).numpy() * 3000.0
    b = S.render(shape, x + 1.3, y - 0.7, inten).numpy() * 3000.0
    a, b = np.round(a), np.round(b)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    got, ok = bench.planes_fp64(torch.tensor(a), torch.tensor(b), wc.win_ctrs_x, wc.win_ctrs_y, n, w)
    ref, ok_ref = O.planes_fp64(a, b, wc.win_ctrs_x.astype(np.float32), wc.win_ctrs_y.astype(np.float32), (n, n),
                                w.astype(np.float64), w.astype(np.float64))
    ok = ok.numpy().reshape(-1)
    np.testing.assert_array_equal(ok, ok_ref)
    g = got.numpy().reshape(-1, n, n)[ok]
    r = ref[ok]
    assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-12
    # the zero-lag bin of the auto-correlation is the window energy
    auto, _ = bench.planes_fp64(torch.tensor(a), torch.tensor(a), wc.win_ctrs_x, wc.win_ctrs_y, n, w)
    k = int(np.flatnonzero(ok)[0])
    wa, _, _ = O.windows_fp64(a, a, wc.win_ctrs_x.astype(np.float32), wc.win_ctrs_y.astype(np.float32), (n, n),
                              w.astype(np.float64), w.astype(np.float64))
    e = float((wa[k] ** 2).sum())
    assert abs(auto.reshape(-1, n, n)[k, n // 2, n // 2].item() - e) < 1e-9 * e

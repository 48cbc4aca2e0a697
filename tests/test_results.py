"""Per-pair results and the reference's output files (no GPU: the pass
results are built from CPU tensors)."""

import json

import numpy as np
import pytest
import scipy.io
import torch

from pivtools_amd.multipass import PassResult
from pivtools_amd.results import (FULL_FIELDS, MINIMAL_FIELDS, pair_result, save_corr_plane_dump,
                                  save_piv_result, to_half)


def _passes(dump=False, P=2):
    rng = np.random.default_rng(0)
    out = []
    for p, (ny, nx, win) in enumerate([(5, 7, (32, 32)), (9, 13, (16, 16))]):
        N = 3
        u = torch.tensor(rng.normal(0, 2, (2, N, ny, nx)).astype(np.float32))
        u[0, 1, 2, 3] = float("nan")
        u[1, 0, 0, 0] = 1e6                                # beyond float16
        nm = torch.zeros((N, ny, nx), dtype=torch.uint8)
        nm[1, 2, 3] = 1
        rs = torch.zeros((N, ny, nx), dtype=torch.int8)
        rs[1, 2, 3] = 10
        pred = None if p == 0 else torch.tensor(rng.normal(0, 1, (2, N, ny, nx)).astype(np.float32))
        dd = None
        if dump:
            dd = {"planes": torch.tensor(rng.random((N, ny, nx) + win, dtype=np.float32)),
                  "pk_loc_x_raw": torch.zeros((N, P, ny, nx)), "pk_loc_y_raw": torch.zeros((N, P, ny, nx)),
                  "pk_height_raw": torch.ones((N, P, ny, nx)), "pk_ratio_raw": torch.ones((N, ny, nx)),
                  "sx": torch.ones((N, P, ny, nx)), "sy": torch.ones((N, P, ny, nx)),
                  "sxy": torch.zeros((N, P, ny, nx)), "peak_finder": 3}
        out.append(PassResult(ux=u[1], uy=u[0], nan_mask=nm, nan_reason=rs,
                              peak_mag=torch.rand((N, P, ny, nx)), peak_ratio=torch.rand((N, ny, nx)) * 5,
                              predictor=pred, window_size=win, n_windows=(ny, nx),
                              primary_peak_mag=torch.rand((N, ny, nx)),
                              peak_choice=torch.ones((N, ny, nx), dtype=torch.uint8),
                              b_mask=np.zeros((ny, nx), bool), win_ctrs_x=np.arange(nx, dtype=np.float32) * 8 + 7.5,
                              win_ctrs_y=np.arange(ny, dtype=np.float32) * 8 + 7.5, debug_dump=dd))
    return out


def test_pair_result_slices_one_pair():
    ps = _passes()
    r = pair_result(ps, 1)
    assert len(r.passes) == 2
    p1 = r.passes[1]
    assert p1.ux_mat.shape == (9, 13) and p1.peak_mag.shape == (2, 9, 13)
    assert p1.predictor_field.shape == (9, 13, 2)
    np.testing.assert_array_equal(p1.predictor_field[..., 0], ps[1].predictor[0, 1].numpy())
    assert np.isnan(r.passes[0].uy_mat[2, 3]) and r.passes[0].nan_mask[2, 3]
    assert (r.passes[0].predictor_field == 0).all()


@pytest.mark.parametrize("mode", ["full", "minimal"])
def test_mat_file_layout(tmp_path, mode):
    r = pair_result(_passes(), 0)
    fn = save_piv_result(r, tmp_path, 7, runs_to_save=[1], save_mode=mode)
    assert fn.endswith("B00007.mat")
    m = scipy.io.loadmat(fn)["piv_result"]
    assert m.shape == (1, 2)
    assert set(m.dtype.names) == set(FULL_FIELDS if mode == "full" else MINIMAL_FIELDS)
    assert m["ux"][0, 0].size == 0                      # pass not in runs_to_save
    ux = m["ux"][0, 1]
    # MAT files have no float16: the half-precision values are stored as doubles
    assert ux.shape == (9, 13)
    np.testing.assert_array_equal(ux, to_half(r.passes[1].ux_mat).astype(np.float64))
    assert ux[0, 0] == 65504.0
    if mode == "full":
        assert m["nan_reason"][0, 1].dtype == np.int8
        assert m["peak_ratio"][0, 1].dtype == np.float32
        assert m["peak_mag"][0, 1].shape == (2, 9, 13)
        assert m["predictor_field"][0, 1].shape == (9, 13, 2)
        np.testing.assert_array_equal(m["window_size"][0, 1].ravel(), [16, 16])
        np.testing.assert_array_equal(m["n_windows"][0, 1].ravel(), [9, 13])


def test_half_precision_clamps_and_keeps_nan():
    a = np.array([1e6, -1e6, np.nan, 1.5, np.inf], np.float32)
    h = to_half(a)
    assert h.dtype == np.float16
    assert h[0] == 65504 and h[1] == -65504 and np.isnan(h[2]) and h[3] == 1.5 and h[4] == 65504


def test_corr_plane_dump(tmp_path):
    assert save_corr_plane_dump(pair_result(_passes(), 0), tmp_path, 1) is None
    ps = _passes(dump=True)
    r = pair_result(ps, 2)
    fn = save_corr_plane_dump(r, tmp_path, 1)
    assert fn.endswith("debug_corr_planes/B00001_corrplanes.npz")
    with np.load(fn, allow_pickle=False) as z:
        assert int(z["frame_number"]) == 1 and int(z["n_passes"]) == 2
        labels = json.loads(str(z["nan_reason_labels"]))
        assert "-1" in labels and "0" in labels
        np.testing.assert_array_equal(z["pass1_planes"], ps[1].debug_dump["planes"][2].numpy())
        assert z["pass1_planes"].shape == (9, 13, 16, 16)
        np.testing.assert_array_equal(z["pass0_nan_reason"], r.passes[0].nan_reason)
        assert int(z["pass0_peak_finder"]) == 3

"""pivamd_xcorr_noalloc (the correlation entry point that allocates nothing,
for callers that cannot call hipMalloc) against pivamd_xcorr with a caller
workspace: bit-identical outputs on two calls in a row, the second with a
smaller input, through the C ABI.  Also: a call whose peak records exceed
the device scratch is refused with PIVAMD_ERR_SIZE before any launch."""

import numpy as np
import pytest
import torch

from pivtools_amd import _native
from pivtools_amd import synthetic as S
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype)


def _run(fn, a, b, win, n_peaks, finder, with_work):
    L = _native.core()
    N, H, W = a.shape
    wc = compute_window_centers((H, W), (win, win), 50)
    ny, nx = wc.n_win_y, wc.n_win_x
    w = _t(window_weight((win, win), "gaussian"))
    pw = torch.empty((win, win), device=dev)
    st = torch.cuda.current_stream().cuda_stream
    assert L.pivamd_plane_weight(w.data_ptr(), win, win, pw.data_ptr(), st) == 0
    cx, cy = _t(wc.win_ctrs_x), _t(wc.win_ctrs_y)
    mask = torch.zeros((ny, nx), device=dev)
    mask[0, 1] = 1.0
    shape = (N, n_peaks, ny, nx)
    outs = [torch.full(shape, -7.0, device=dev) for _ in range(6)]
    ratio = torch.full((N, ny, nx), -7.0, device=dev)
    args = [a.data_ptr(), b.data_ptr(), mask.data_ptr(), H, W, N, cx.data_ptr(), cy.data_ptr(), ny, nx,
            w.data_ptr(), w.data_ptr(), pw.data_ptr(), 0, win, win, n_peaks, finder, 0,
            *[o.data_ptr() for o in outs], ratio.data_ptr(), None]
    if with_work:
        nb = L.pivamd_xcorr_workspace(N, ny, nx, win, win, n_peaks, 0)
        work = torch.empty(max(nb, 4), dtype=torch.uint8, device=dev)
        rc = fn(*args, work.data_ptr(), nb, st)
    else:
        rc = fn(*args, st)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs] + [ratio.cpu().numpy()]


def _frames(n, shape, seed):
    A, B = [], []
    for k in range(n):
        a, b = S.multipass_pair(seed=seed + k, shape=shape)
        A.append(a.numpy().view(np.uint16).astype(np.float32))
        B.append(b.numpy().view(np.uint16).astype(np.float32))
    return _t(np.stack(A)), _t(np.stack(B))


@pytest.mark.parametrize("win,n_peaks,finder", [(32, 1, 3), (64, 1, 6), (16, 3, 3)])
def test_noalloc_equals_workspace_entry_point_twice(win, n_peaks, finder):
    L = _native.core()
    big = _frames(3, (384, 448), 9100)
    small = _frames(1, (160, 192), 9200)
    for a, b in (big, small):
        print(f"noalloc: win {win} peaks {n_peaks} finder {finder} frames {tuple(a.shape)}", flush=True)
        ref = _run(L.pivamd_xcorr, a, b, win, n_peaks, finder, True)
        got = _run(L.pivamd_xcorr_noalloc, a, b, win, n_peaks, finder, False)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        assert np.isfinite(ref[0][:, 0]).mean() > 0.5
        assert (ref[0][:, :, 0, 1] == -7.0).all()        # the masked window is left untouched


def test_noalloc_refuses_calls_beyond_its_scratch():
    L = _native.core()
    cap = L.pivamd_xcorr_noalloc_max_bytes()
    assert cap == 64 << 20
    # 16-point windows of 64 pairs of 2160 x 2560: 0.70 GB of records
    ny, nx, n = 319, 269, 64
    assert L.pivamd_xcorr_workspace(n, ny, nx, 16, 16, 1, 0) > cap
    st = torch.cuda.current_stream().cuda_stream
    dummy = torch.zeros(16, device=dev)
    p = dummy.data_ptr()
    rc = L.pivamd_xcorr_noalloc(p, p, None, 2160, 2560, n, p, p, ny, nx, p, p, p, 0, 16, 16, 1, 3, 0,
                                p, p, p, p, p, p, p, None, st)
    assert rc == 2

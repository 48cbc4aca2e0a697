"""A frame stack of 4 GiB or more: the correlation kernels read window samples
as buffer loads whose 32-bit offsets cover stacks below 4 GiB, and fall back
to plain loads above.  The same pairs correlated as one 4.3 GB stack (plain
loads) and as two halves (buffer loads) give the same bits."""

import pytest
import torch

from pivtools_amd import _native
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _xcorr(a, b, win):
    L = _native.core()
    N, H, W = a.shape
    wc = compute_window_centers((H, W), (win, win), 50)
    ny, nx = wc.n_win_y, wc.n_win_x
    w = torch.as_tensor(window_weight((win, win), "gaussian"), device=dev)
    pw = torch.empty((win, win), device=dev)
    st = torch.cuda.current_stream().cuda_stream
    assert L.pivamd_plane_weight(w.data_ptr(), win, win, pw.data_ptr(), st) == 0
    cx = torch.as_tensor(wc.win_ctrs_x, device=dev)
    cy = torch.as_tensor(wc.win_ctrs_y, device=dev)
    outs = [torch.zeros((N, 1, ny, nx), device=dev) for _ in range(6)]
    ratio = torch.zeros((N, ny, nx), device=dev)
    nb = L.pivamd_xcorr_workspace(N, ny, nx, win, win, 1, 0)
    work = torch.empty(max(nb, 4), dtype=torch.uint8, device=dev)
    rc = L.pivamd_xcorr(a.data_ptr(), b.data_ptr(), None, H, W, N, cx.data_ptr(), cy.data_ptr(), ny, nx,
                        w.data_ptr(), w.data_ptr(), pw.data_ptr(), 0, win, win, 1, 3, 0,
                        *[o.data_ptr() for o in outs], ratio.data_ptr(), None, work.data_ptr(), nb, st)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return outs + [ratio]


@pytest.mark.parametrize("win", [16, 64])
def test_stack_past_4_gib_equals_its_halves(win):
    H, W, N = 2160, 2560, 196
    assert N * H * W * 4 >= 1 << 32
    g = torch.Generator(device=dev).manual_seed(11)
    print(f"large stack: {N} x {H} x {W} float32, window {win}", flush=True)
    a = torch.rand((N, H, W), generator=g, device=dev)
    b = torch.roll(a, shifts=(1, -2), dims=(1, 2)).contiguous()
    whole = _xcorr(a, b, win)
    half = N // 2
    first = _xcorr(a[:half].contiguous(), b[:half].contiguous(), win)
    second = _xcorr(a[half:].contiguous(), b[half:].contiguous(), win)
    for w_, f_, s_ in zip(whole, first, second):
        assert torch.equal(w_.nan_to_num(7.0), torch.cat([f_, s_]).nan_to_num(7.0))
    # the shift was found: B is A moved by (dy, dx) = (1, -2)
    ok = torch.isfinite(whole[0]) & torch.isfinite(whole[1])
    assert ok.float().mean() > 0.9
    assert round(float(whole[0][ok].median())) == -2 and round(float(whole[1][ok].median())) == 1

"""The Poiseuille pair of unit-tests/generate_poiseuille_ensemble.py restated:
500 x 500, 30000 particles of diameter 3 px (sigma = d / 2.355), peak
intensity U[200, 255], u_x(y) = 5 (1 - (2y/H - 1)^2), u_y = 0, both frames
normalised to the uint16 range by the pair's own maximum; and the error
statistics of unit-tests/test_multipass_convergence.py:80-100,398-430."""

import numpy as np

from pivtools_amd import synthetic as S

SHAPE = (500, 500)
U_MAX = 5.0
SCHEDULE = [(64, 64), (32, 32), (16, 16)]
OVERLAP = [50, 50, 50]
# the reference test's configuration: gauss6, median 5x5 threshold 3 eps 0.1,
# nearest infill mid-pass, biharmonic on the final pass
OUTLIERS = [{"type": "median_2d", "threshold": 3.0, "epsilon": 0.1}]


def pair(seed=101):
    rng = np.random.default_rng(seed)
    H, W = SHAPE
    x = rng.uniform(0, W, 30_000)
    y = rng.uniform(0, H, 30_000)
    inten = rng.uniform(200, 255, 30_000)
    dx = U_MAX * (1.0 - (2.0 * y / H - 1.0) ** 2)
    sigma = 3.0 / 2.355
    a = S.render(SHAPE, x, y, inten, sigma=sigma).numpy().astype(np.float64)
    b = S.render(SHAPE, x + dx, y, inten, sigma=sigma).numpy().astype(np.float64)
    out = []
    for img in (a, b):
        out.append((img / img.max() * 65535).astype(np.uint16))
    return out[0], out[1]


def errors(ux, uy, valid):
    """Median |ux - Poiseuille| and median |uy| over the interior rows (15 %
    of the rows excluded at each wall), on the reference's row positions
    np.linspace(0, H - 1, nrows)."""
    nrows = ux.shape[0]
    margin = max(1, int(nrows * 0.15))
    interior = np.zeros(ux.shape, bool)
    interior[margin:-margin, :] = True
    interior &= valid
    rows = np.linspace(0, SHAPE[0] - 1, nrows)
    gt = U_MAX * (1.0 - (2.0 * rows / SHAPE[0] - 1.0) ** 2)
    gt = np.broadcast_to(gt[:, None], ux.shape)
    return float(np.median(np.abs(ux[interior] - gt[interior]))), float(np.median(np.abs(uy[interior])))

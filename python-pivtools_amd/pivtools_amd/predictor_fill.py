"""Masked windows of a predictor field: the harmonic fill as a linear map.

Masked windows hold 0, which is not data; before a pass's field becomes the
next pass's predictor the reference replaces them with a fill from the
unmasked windows (pivtools_cli/piv/piv_backend/cpu_instantaneous.py:843-854,
infilling.py:808-870 ``fill_masked_predictor``, default 'harmonic').  The
harmonic fill (infilling.py:144-260, order 1, no excluded nodes) minimises
the sum of squared first differences, in physical units, over the difference
equations that touch a masked node, with every unmasked node held at its
value.  Its normal equations read, for each masked node a,

    sum_b w_ab (u_a - u_b) = 0,   w = 1/py^2 along y, 1/px^2 along x,

over a's in-grid neighbours b.  With the masked unknowns x and the unmasked
neighbours u_B of the mask this is A x = R u_B, so x = (A^-1 R) u_B: a fixed
linear map that depends on the mask and the grid pitch only.

``harmonic_operator`` solves for that map once per pass on the host (sparse
LU, FP64, like the reference's solve) and the device applies it to every
field of every batch (``pivamd_fill_linear``: one FP64 dot product per
masked window and field).  The reference's errors are kept: a group of
masked windows with no unmasked neighbour cannot be filled.
"""

from __future__ import annotations

import numpy as np

# the map is dense (masked windows x unmasked neighbours); above this many
# entries (512 MB of FP64) the 'nearest' fill is the one to use
MAX_ENTRIES = 1 << 26


def harmonic_operator(mask: np.ndarray, spacing, max_entries: int = MAX_ENTRIES):
    """Return (fill_idx int32 (nf,), bnd_idx int32 (nb,), gt float64 (nb, nf))
    with x[fill_idx] = gt.T @ u[bnd_idx] for a field u on the (ny, nx) grid."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    from scipy.sparse.linalg import splu

    m = np.asarray(mask, bool)
    ny, nx = m.shape
    py, px = (float(s) for s in spacing)
    if not (np.isfinite(py) and np.isfinite(px) and py > 0 and px > 0):
        raise ValueError(f"spacing must be positive, got {spacing!r}")
    fill_idx = np.flatnonzero(m.ravel())
    nf = fill_idx.size
    if nf == 0:
        return fill_idx.astype(np.int32), np.zeros(0, np.int32), np.zeros((0, 0))
    if nf == m.size:
        raise ValueError("every window is masked: the harmonic fill has no valid node")
    pos = np.full(m.size, -1, np.int64)
    pos[fill_idx] = np.arange(nf)
    idx = np.arange(m.size).reshape(ny, nx)
    ar, ac, av = [], [], []          # A (nf x nf)
    rr, rc, rv = [], [], []          # R (nf x grid nodes)
    mf = m.ravel()
    for a, b, h in ((idx[:-1, :], idx[1:, :], py), (idx[:, :-1], idx[:, 1:], px)):
        a, b = a.ravel(), b.ravel()
        w = 1.0 / (h * h)
        # each neighbour pair is one difference equation; it enters the
        # normal equations of each of its masked ends
        for s_, t_ in ((a, b), (b, a)):
            fs, ft = mf[s_], mf[t_]
            ar.append(pos[s_[fs]])
            ac.append(pos[s_[fs]])
            av.append(np.full(int(fs.sum()), w))
            cpl = fs & ft                         # both ends masked: coupling
            ar.append(pos[s_[cpl]])
            ac.append(pos[t_[cpl]])
            av.append(np.full(int(cpl.sum()), -w))
            kn = fs & ~ft                         # masked s next to unmasked t
            rr.append(pos[s_[kn]])
            rc.append(t_[kn])
            rv.append(np.full(int(kn.sum()), w))
    A = sp.csr_matrix((np.concatenate(av), (np.concatenate(ar), np.concatenate(ac))), shape=(nf, nf))
    R = sp.csr_matrix((np.concatenate(rv), (np.concatenate(rr), np.concatenate(rc))), shape=(nf, m.size))
    bnd_idx = np.unique(R.indices)
    nb = bnd_idx.size
    n_groups, group = connected_components(A, directed=False)
    anchored = np.zeros(n_groups, bool)
    anchored[np.unique(group[np.flatnonzero(np.diff(R.indptr) > 0)])] = True
    if not anchored.all():
        raise ValueError(f"{int((~anchored).sum())} group(s) of masked windows have no unmasked neighbour; "
                         "the harmonic fill cannot determine them")
    if nf * nb > max_entries:
        raise ValueError(f"the harmonic fill map of this mask has {nf} x {nb} entries (limit {max_entries}); "
                         "use the 'nearest' predictor fill")
    Rb = R[:, bnd_idx].toarray()
    G = splu(A.tocsc()).solve(Rb)                 # (nf, nb)
    resid = np.linalg.norm(A @ G - Rb) / max(np.linalg.norm(Rb), np.finfo(float).tiny)
    if not np.all(np.isfinite(G)) or resid > 1e-6:
        raise np.linalg.LinAlgError(f"numerically singular harmonic fill system (relative residual {resid:.2e})")
    return fill_idx.astype(np.int32), bnd_idx.astype(np.int32), np.ascontiguousarray(G.T)

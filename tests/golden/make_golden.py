"""Write golden vectors from the reference's own C sources (oracle/_ref).

Run in a container where the reference is mounted and ``make -C oracle``
has built oracle/_ref/libref_peak.so and libref_warp.so.  The fixtures are
data (inputs and the reference's outputs); the tests compare the oracle and
the device path against them without needing the reference.

    python tests/golden/make_golden.py
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def peak_vectors(rng, count=192):
    """Random Gaussian-plus-noise planes of several sizes, fitted by the
    reference's lsqpeaklocate_lm (scalar) and lsqpeaklocate_lm_batch."""
    lib = ctypes.CDLL(os.path.join(REF, "libref_peak.so"))
    lib.lsqpeaklocate_lm.argtypes = [f32p, i32p, f32p, ctypes.c_int, ctypes.c_int, f32p, ctypes.c_void_p]
    lib.peak_detectability.argtypes = [f32p, i32p]
    lib.peak_detectability.restype = ctypes.c_float
    lib.lsqpeaklocate_lm_batch.argtypes = [f32p, ctypes.c_int, i32p, ctypes.c_int, f32p, f32p, ctypes.c_void_p]
    lib.peakfit_batch_lanes.restype = ctypes.c_int
    lanes = lib.peakfit_batch_lanes()
    out = {}
    for n in (16, 32, 64):
        planes = []
        for _ in range(count // 3):
            ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            c = n / 2
            i0, j0 = rng.uniform(-n / 5, n / 5, 2)
            s = rng.uniform(0.8, 2.5)
            amp = rng.uniform(0.05, 1.0)
            pl = amp * np.exp(-((ii - c - i0) ** 2 + (jj - c - j0) ** 2) / s ** 2)
            pl += rng.normal(0, rng.uniform(0, 0.08) * amp, (n, n))
            planes.append(pl.astype(np.float32))
        planes = np.array(planes, np.float32)
        w = (1.0 / np.outer(1 - np.abs(np.arange(n) - n / 2) / n, 1 - np.abs(np.arange(n) - n / 2) / n)).astype(np.float32)
        N = np.array([n, n], np.int32)
        for ft in (3, 4, 5, 6):
            loc = np.zeros((len(planes), 3), np.float32)
            std = np.zeros((len(planes), 3), np.float32)
            for k, pl in enumerate(planes):
                lib.lsqpeaklocate_lm(np.ascontiguousarray(pl), N, loc[k], 1, ft, std[k], w.ctypes.data)
            out[f"n{n}_ft{ft}_loc"] = loc
            out[f"n{n}_ft{ft}_std"] = std
            if ft >= 4:
                bl = np.zeros((len(planes), 3), np.float32)
                bs = np.zeros((len(planes), 3), np.float32)
                for s0 in range(0, len(planes), lanes):
                    chunk = np.ascontiguousarray(planes[s0:s0 + lanes])
                    L = len(chunk)
                    pl_loc = np.zeros(3 * lanes, np.float32)
                    pl_std = np.zeros(3 * lanes, np.float32)
                    lib.lsqpeaklocate_lm_batch(chunk.reshape(-1), L, N, ft, pl_loc, pl_std, w.ctypes.data)
                    bl[s0:s0 + L] = pl_loc.reshape(3, lanes)[:, :L].T
                    bs[s0:s0 + L] = pl_std.reshape(3, lanes)[:, :L].T
                out[f"n{n}_ft{ft}_batch_loc"] = bl
                out[f"n{n}_ft{ft}_batch_std"] = bs
        out[f"n{n}_planes"] = planes
        out[f"n{n}_weight"] = w
        out[f"n{n}_ratio"] = np.array([lib.peak_detectability(np.ascontiguousarray(p), N) for p in planes],
                                      np.float32)
    return out


def warp_vectors(rng):
    lib = ctypes.CDLL(os.path.join(REF, "libref_warp.so"))
    lib.fused_symmetric_warp_batch.argtypes = [f32p] * 6 + [ctypes.c_int] * 5 + [f32p, f32p] + [ctypes.c_int] * 3
    lib.fused_warp_set_impl.argtypes = [ctypes.c_int]
    lib.fused_warp_set_impl(0)
    N, H, W = 2, 64, 80
    A = rng.uniform(0, 4095, (N, H, W)).astype(np.float32)
    B = rng.uniform(0, 4095, (N, H, W)).astype(np.float32)
    cy = np.array([-10, 7.5, 23.5, 39.5, 55.5, 70], np.float32)
    cx = np.array([-10, 7.5, 23.5, 39.5, 55.5, 71.5, 90], np.float32)
    dy = rng.uniform(-3, 3, (N, len(cy), len(cx))).astype(np.float32)
    dx = rng.uniform(-3, 3, (N, len(cy), len(cx))).astype(np.float32)
    out = {"A": A, "B": B, "cy": cy, "cx": cx, "dy": dy, "dx": dx}
    for mode in (0, 1):
        for shared in (0, 1):
            for rnd in (0, 1):
                oa = np.zeros_like(A)
                ob = np.zeros_like(B)
                pdy = np.ascontiguousarray(dy[0] if shared else dy)
                pdx = np.ascontiguousarray(dx[0] if shared else dx)
                r = lib.fused_symmetric_warp_batch(A, B, oa, ob, pdy, pdx, N, H, W, len(cy), len(cx), cy, cx,
                                                   mode, shared, rnd)
                assert r == 0
                out[f"m{mode}_s{shared}_r{rnd}_a"] = oa
                out[f"m{mode}_s{shared}_r{rnd}_b"] = ob
    return out


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    np.savez_compressed(os.path.join(HERE, "peakfit_ref.npz"), **peak_vectors(rng))
    np.savez_compressed(os.path.join(HERE, "warp_ref.npz"), **warp_vectors(rng))
    print("wrote", os.listdir(HERE))

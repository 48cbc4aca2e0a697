"""Extract the camera-payload vectors used by the ingest tests.

Source: the reference's `.pivraw` contract fixtures
(unit-tests/fixtures/pivraw_contract/, written by the acquisition app's own
recorder) and their decoded frames (`expected_arrays.npz` there).  This
script reads those files as data only -- the JSON sidecar for the geometry,
`numpy.fromfile` for the frame payloads, `numpy.load` (no pickle) for the
expected frames -- and imports nothing from the reference.  It keeps each
frame's payload bytes (not the 4096-byte stride padding) so the fixture stays
small; the tests place the payloads back at a stride themselves.

Run from the repository root:  python tests/golden/make_pivraw_vectors.py
"""

from __future__ import annotations

import json
import os

import numpy as np

SRC = "/root/reference/unit-tests/fixtures/pivraw_contract"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pivraw_vectors.npz")

# (case, camera, has_dark)
CASES = [
    ("df_2cam_clean", 1, False),
    ("df_2cam_clean", 2, False),
    ("df_2cam_dark", 1, True),
    ("df_2cam_dark", 2, True),
    ("tr_bayer_clean", 1, False),
    ("tr_bayer_dark", 1, True),
]


def _payloads(path: str):
    with open(path + ".json") as f:
        hdr = json.load(f)
    n, bpf, stride = int(hdr["n_frames"]), int(hdr["bytes_per_frame"]), int(hdr["frame_stride"])
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < (n - 1) * stride + bpf:
        raise ValueError(f"{path}: {raw.size} bytes for {n} frames of stride {stride}")
    frames = np.stack([raw[k * stride:k * stride + bpf] for k in range(n)])
    meta = {"width": int(hdr["width"]), "height": int(hdr["height"]), "pixel_format": hdr["pixel_format"],
            "frame_stride": stride, "bytes_per_frame": bpf, "mode": hdr["mode"], "n_frames": n}
    return frames, meta


def main() -> None:
    out = {}
    meta_all = {}
    with np.load(os.path.join(SRC, "expected_arrays.npz"), allow_pickle=False) as exp:
        for case, cam, has_dark in CASES:
            key = f"{case}__Camera{cam}"
            frames, meta = _payloads(os.path.join(SRC, case, f"Camera{cam}.pivraw"))
            out[key + "__payload"] = frames
            out[key + "__decoded"] = exp[key]
            if has_dark:
                dframes, dmeta = _payloads(os.path.join(SRC, case, "dark", f"Camera{cam}.pivraw"))
                out[key + "__dark_payload"] = dframes
                out[key + "__dark_subtracted"] = exp[key + "__dark_subtracted"]
                meta["dark_mode"] = dmeta["mode"]
                meta["dark_frames"] = dmeta["n_frames"]
            meta_all[key] = meta
    out["meta_json"] = np.frombuffer(json.dumps(meta_all, sort_keys=True).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {sorted(k for k in out if k != 'meta_json')}")


if __name__ == "__main__":
    main()

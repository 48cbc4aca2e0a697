"""Collect the DaVis container vectors used by the reader tests.

Source: the reference's reader fixtures, each LaVision's own answer (lvpyio)
for small synthetic recordings, read here as data only -- raw file bytes and
``numpy.load`` (no pickle) -- with nothing imported from the reference:

  unit-tests/fixtures/lavision_transformers/lvpyio_oracle.npz
      whole .set recordings (every file's bytes) for all geometries, all three
      file-backed corrections and their orders, with lvpyio's frames; copied
      unchanged;
  unit-tests/fixtures/set_packed/{id}.payload.npy, {id}.lvpyio.npy
      one 64 x 48 payload per packed encoding and lvpyio's frame;
  unit-tests/fixtures/im7_pack1/*.im7, im7_v1_pack20/*.im7 (+ .lvpyio.npy)
      .im7 files with the token stream (pack 1) and per-frame LZ4 (pack 20);
  unit-tests/fixtures/set_imx/fixture.set, fixture/*, fixture.lvpyio.npy
      an IMX-compressed mono-12p recording (107 x 5312) and lvpyio's frame.

Run from the repository root:  python tests/golden/make_davis_vectors.py
"""

from __future__ import annotations

import os
import shutil

import numpy as np

SRC = "/root/reference/unit-tests/fixtures"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "davis")


def _file(path: str) -> np.ndarray:
    return np.fromfile(path, dtype=np.uint8)


def main() -> None:
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(SRC, "lavision_transformers", "lvpyio_oracle.npz"),
                    os.path.join(OUT, "transformers_lvpyio.npz"))
    packed = {}
    for enc in ("mono-10pmsb", "mono-12pmsb", "mono-12packed"):
        key = enc.replace("-", "_")
        packed[key + "__payload"] = np.load(os.path.join(SRC, "set_packed", f"{enc}.payload.npy"))
        packed[key + "__lvpyio"] = np.load(os.path.join(SRC, "set_packed", f"{enc}.lvpyio.npy"))
    np.savez_compressed(os.path.join(OUT, "packed_lvpyio.npz"), **packed)
    im7 = {}
    for sub, name in (("im7_pack1", "carry_3frames"), ("im7_pack1", "high_nibble_8"),
                      ("im7_v1_pack20", "three_frames")):
        im7[f"{name}__im7"] = _file(os.path.join(SRC, sub, f"{name}.im7"))
        im7[f"{name}__lvpyio"] = np.load(os.path.join(SRC, sub, f"{name}.lvpyio.npy"))
    np.savez_compressed(os.path.join(OUT, "im7_lvpyio.npz"), **im7)
    imx = {"fixture.set": _file(os.path.join(SRC, "set_imx", "fixture.set"))}
    for f in sorted(os.listdir(os.path.join(SRC, "set_imx", "fixture"))):
        imx["fixture~" + f] = _file(os.path.join(SRC, "set_imx", "fixture", f))
    imx["lvpyio"] = np.load(os.path.join(SRC, "set_imx", "fixture.lvpyio.npy"))
    np.savez_compressed(os.path.join(OUT, "set_imx_lvpyio.npz"), **imx)
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()

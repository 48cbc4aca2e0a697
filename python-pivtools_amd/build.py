"""Build the gfx950 libraries in-tree (python-pivtools_amd/pivtools_amd/lib/).

    libpivtools_amd.so   device kernels + the device-pointer C ABI (pivamd_*)
    libbulkxcorr2d.so    host-pointer drop-in for the reference's libbulkxcorr2d
    libfusedwarp.so      host-pointer drop-in for the reference's libfusedwarp
    libkspacefit.so      host-pointer drop-in for the reference's libkspacefit

Run ``python python-pivtools_amd/build.py`` (or ``__graft_entry__.build()``).
Translation units compile in parallel; every HIP line carries the
packed-FP32 opt-out.
"""

from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "pivtools_amd", "lib")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("PIVAMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

# Kernel translation units of libpivtools_amd.so.  xcorr.hip is split by
# window size so the expensive template instantiations build in parallel.
XCORR_SIZES = (8, 12, 16, 24, 32, 48, 64, 96, 128)
KERNEL_TUS = [("warp", "warp.hip", [])]
KERNEL_TUS += [("xcorr_common", "xcorr.hip", ["-DPIVAMD_XCORR_COMMON"])]
KERNEL_TUS += [(f"xcorr_{n}", "xcorr.hip", [f"-DPIVAMD_XCORR_SIZE={n}"]) for n in XCORR_SIZES]
EXTRA_TUS = ["ensemble.hip", "glue.hip", "infill.hip", "ingest.hip", "kspace.hip", "outliers.hip", "readers.cpp",
             "jpeg.cpp"]


def _run_parallel(jobs):
    procs = []
    for name, cmd in jobs:
        procs.append((name, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for name, cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
        if p.returncode != 0:
            failed.append((name, " ".join(cmd), text))
        elif text.strip():
            print(f"[{name}]\n{text}")
    if failed:
        for name, cmd, text in failed:
            print(f"FAILED {name}: {cmd}\n{text}", file=sys.stderr)
        raise RuntimeError(f"build failed: {[f[0] for f in failed]}")


def build(verbose: bool = True) -> None:
    t0 = time.time()
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    objs = []
    tus = list(KERNEL_TUS) + [(os.path.splitext(f)[0], f, []) for f in EXTRA_TUS
                             if os.path.isfile(os.path.join(CSRC, f))]
    for name, src, defs in tus:
        obj = os.path.join(OBJ, name + ".o")
        objs.append(obj)
        jobs.append((name, [HIPCC, *COMMON, *defs, "-c", os.path.join(CSRC, src), "-o", obj]))
    for flavour in ("XCORR", "WARP", "KSPACE"):
        obj = os.path.join(OBJ, f"capi_{flavour.lower()}.o")
        jobs.append((f"capi_{flavour.lower()}",
                     [HIPCC, *COMMON, f"-DPIVAMD_CAPI_{flavour}", "-c", os.path.join(CSRC, "capi.cpp"), "-o", obj]))
    n_par = int(os.environ.get("MAX_JOBS", "8"))
    for i in range(0, len(jobs), n_par):
        _run_parallel(jobs[i:i + n_par])
    core = os.path.join(OUT, "libpivtools_amd.so")
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", core + ".tmp", *objs]
    subprocess.run(link, check=True)
    os.replace(core + ".tmp", core)
    for flavour, lib in (("xcorr", "libbulkxcorr2d.so"), ("warp", "libfusedwarp.so"),
                         ("kspace", "libkspacefit.so")):
        path = os.path.join(OUT, lib)
        cmd = [HIPCC, "-shared", "-fPIC", "-o", path + ".tmp", os.path.join(OBJ, f"capi_{flavour}.o"),
               f"-L{OUT}", "-lpivtools_amd", "-Wl,-rpath,$ORIGIN"]
        subprocess.run(cmd, check=True)
        os.replace(path + ".tmp", path)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    build()

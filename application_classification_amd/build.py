"""Build recipe for the native parts of the package (run by ``__graft_entry__.build()``).

Produces, in this directory:
  * ``libac_mi355x.so`` — the C-ABI library (include/ac_mi355x.h): HIP kernels for gfx950 + host runtime.
  * ``ac_main``         — the command-line driver with the reference's CLI (``ac_main data.bin pattern.bin``).
Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libac_mi355x.so")
EXE = os.path.join(HERE, "ac_main")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIP_SOURCES = ["ac_kernels.hip", "ac_stage_ops.hip", "ac_adjoint.hip", "ac_capi.hip"]
HOST_SOURCES = ["ac_gen.cpp", "ac_io.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = [ln for ln in (r.stderr or "").splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False, jobs: int = 8) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJ, exist_ok=True)
    steps = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        steps.append([HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        steps.append([HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), steps))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread", "-L/opt/rocm/lib",
          "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "ac_main.cpp")
    if os.path.exists(main_src):
        tmpx = EXE + ".tmp"
        _run([HIPCC, *COMMON, main_src, "-o", tmpx, f"-L{HERE}", "-lac_mi355x", "-Wl,-rpath,$ORIGIN", "-pthread"],
             verbose)
        os.replace(tmpx, EXE)
    return LIB


if __name__ == "__main__":
    build(verbose=True)

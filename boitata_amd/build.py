"""Builds libboitata_gpu.so, the HIP library behind include/boitata_gpu.h.

hipcc for gfx950 only. Every translation unit is compiled with
-ffp-contract=off (the FP64 path reproduces the reference's unfused
arithmetic; FP32 code opts into contraction locally) and with packed FP32
instructions disabled. Run as ``python -m boitata_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libboitata_gpu.so"
# (source, object, extra flags). bk_solve.hip is built as two objects: the
# kriging pass and the persistent refined gradient with their loops over work
# units (BK_PART 1), where machine LICM would hoist FP64 constants (polynomial
# coefficients, zeroed accumulators) out of the loop and spill them, so it is
# off there; and the other refined gradients (BK_PART 2), which have no such
# loop and keep it.
UNITS = [
    ("bk_search.hip", "bk_search.o", []),
    ("bk_solve.hip", "bk_solve_pass.o", ["-DBK_PART=1", "-mllvm", "-disable-machine-licm"]),
    ("bk_solve.hip", "bk_solve_grad.o", ["-DBK_PART=2"]),
    ("bk_mik.hip", "bk_mik.o", []),
    ("bk_diag.hip", "bk_diag.o", []),
    ("bk_api.hip", "bk_api.o", []),
]
SOURCES = sorted({u[0] for u in UNITS})
ARCH = os.environ.get("BK_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> Path:
    """Compiles every source (the units in parallel) and links the shared
    library; returns its path."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out_dir = CSRC / "_obj"
    out_dir.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for src, name, extra in UNITS:
        obj = out_dir / name
        cmds.append([hipcc, *FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)

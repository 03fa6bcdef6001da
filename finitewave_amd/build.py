"""Builds the gfx950 shared library ``finitewave_amd/_lib/libfinitewave_amd.so``.

One translation unit per (precision, model) pair for the fused step kernel,
compiled in parallel: FP64 units with ``-ffp-contract=off`` (no fused
multiply-add, the reference's arithmetic), FP32 units with ``-ffp-contract=on``:
multiply-adds fused within one source expression only, so every kernel that
inlines a formula rounds it the same way (the backend's cross-statement
fusion under ``fast`` depends on each kernel's basic blocks).
Packed FP32 instructions are disabled in every unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libfinitewave_amd.so"
ARCH = os.environ.get("FW_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", str(CSRC)]
FP64 = ["-ffp-contract=off"]
FP32 = ["-ffp-contract=on", "-fno-math-errno", "-freciprocal-math", "-fapprox-func",
        "-fno-signed-zeros",
        "-fno-trapping-math"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _units(build_dir: Path):
    units = [(CSRC / "fw_abi.hip", build_dir / "fw_abi.o", FP64, []),
             (CSRC / "fw_literal.hip", build_dir / "fw_literal.o", FP64, []),
             (CSRC / "fw_track.hip", build_dir / "fw_track.o", FP64, [])]
    for prec in (0, 1):
        for model in range(8):
            for stem in ("fw_step_inst", "fw_batch_inst"):
                units.append((CSRC / f"{stem}.hip", build_dir / f"{stem}_p{prec}_m{model}.o",
                              FP64 if prec == 0 else FP32,
                              [f"-DFW_PREC={prec}", f"-DFW_MODEL={model}"]))
    return units


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    hipcc = _hipcc()
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    OUT_DIR.mkdir(exist_ok=True)
    units = _units(build_dir)
    jobs = jobs or min(8, os.cpu_count() or 4)

    def compile_one(u):
        src, obj, flags, defs = u
        cmd = [hipcc, "-c", str(src), "-o", str(obj), *COMMON, *flags, *defs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print("built", obj.name, flush=True)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, units))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
           *[str(o) for o in objs]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    import time
    t0 = time.time()
    out = build(verbose=True)
    print(f"{out} ({time.time() - t0:.1f}s)")
    sys.exit(0)

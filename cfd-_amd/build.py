"""Builds the gfx950 shared library of the package: cfd2d_amd/libcfd2d_hip.so.

    python cfd-_amd/build.py          (hipcc cross-compiles; no GPU needed)

The whole library is one translation unit (csrc/capi.hip includes the kernel
files).  Flags:
  -ffp-contract=off   keep every multiply and add separately rounded, so the
                      stencils reproduce the reference's double arithmetic;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      never emit packed FP32 (v_pk_*_f32) instructions.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc" / "capi.hip"
OUT = HERE / "cfd2d_amd" / "libcfd2d_hip.so"
ARCH = os.environ.get("CFD2D_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: Path = OUT) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-shared", "-fPIC",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", str(HERE / "csrc"), "-o", str(out), str(SRC)]


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

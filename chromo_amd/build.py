"""Build recipe for libchromo_hip.so (gfx950).

    python -m chromo_amd.build

compiles chromo_amd/csrc/chromo_hip.hip and chromo_energy.hip with hipcc into chromo_amd/libchromo_hip.so.
-ffp-contract=off keeps every multiply and add separately rounded, as in the
reference's x86-64 build; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SRCS = [HERE / "csrc" / "chromo_hip.hip", HERE / "csrc" / "chromo_energy.hip"]
OUT = HERE / "libchromo_hip.so"
ARCH = os.environ.get("CHROMO_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC", "-shared", f"-I{ROOT / 'include'}", f"-I{HERE / 'csrc'}",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found")
    return p


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *[str(s) for s in SRCS]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

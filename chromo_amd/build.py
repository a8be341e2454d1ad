"""Build recipe for libchromo_hip.so (gfx950).

    python -m chromo_amd.build

Compiles the translation units below with hipcc, in parallel, and links them into
chromo_amd/libchromo_hip.so:

* chromo_hip.hip: one-wavefront kernels, RNG, density, tape, host entry points;
* chromo_energy.hip: energies and their adjoint;
* chromo_spec.hip, once per W in SPEC_WAVES (2): the speculative kernel with W
  wavefronts per replica.

-ffp-contract=off keeps every multiply and add separately rounded, as in the
reference's x86-64 build; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OUT = HERE / "libchromo_hip.so"
ARCH = os.environ.get("CHROMO_OFFLOAD_ARCH", "gfx950")
SPEC_WAVES = (2,)

# (source, extra flags, object name)
UNITS = [(CSRC / "chromo_hip.hip", [], "chromo_hip.o"),
         (CSRC / "chromo_energy.hip", [], "chromo_energy.o")] + \
        [(CSRC / "chromo_spec.hip", [f"-DCHROMO_SPEC_W={w}"], f"chromo_spec_w{w}.o") for w in SPEC_WAVES]

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found")
    return p


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}): {' '.join(cmd[-3:])}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    cc = hipcc()
    jobs = jobs or min(len(UNITS), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with tempfile.TemporaryDirectory(prefix="chromo_build_") as td:
        objs = [str(Path(td) / o) for _, _, o in UNITS]
        cmds = [[cc, *FLAGS, *extra, "-c", str(src), "-o", obj]
                for (src, extra, _), obj in zip(UNITS, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        tmp = OUT.with_suffix(".so.tmp")
        _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs], verbose)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""Builds the CPU oracle (test infrastructure only): oracle/liboracle.so."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return HERE / "liboracle.so"


if __name__ == "__main__":
    print(build())

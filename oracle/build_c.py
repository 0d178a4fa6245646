"""Build the oracle's C restatement into oracle/_bin/ (TEST INFRASTRUCTURE ONLY).

`gcc -O3 -march=x86-64-v2 -fopenmp -shared` on oracle/gs_ref.c. The output is
git-ignored but travels to the GPU host with the tree, like the HIP library.
(The reference itself is Python/JAX, so there is no reference source to
compile into oracle/_ref/.)
"""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_bin" / "libgs_ref.so"


def build() -> Path:
    OUT.parent.mkdir(exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared",
           str(HERE / "gs_ref.c"), "-o", str(OUT), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return OUT


if __name__ == "__main__":
    print(build())

"""Compile the C oracle (test infrastructure) into oracle/liboracle.so with gcc."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "eac_oracle.c"
LIB = HERE / "liboracle.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    # OpenMP only for the all-core baseline (orc_boruvka_omp); every checker
    # function is single-threaded
    # -ffp-contract=off: every product and sum rounded on its own, as the
    # reference's arithmetic is
    cmd = ["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o",
           str(tmp), str(SRC), "-lm"]
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())

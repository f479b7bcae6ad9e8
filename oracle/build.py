"""Compile the C oracle (test infrastructure) into oracle/liboracle.so with gcc."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "eac_oracle.c"
LIB = HERE / "liboracle.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-o", str(tmp), str(SRC), "-lm"]
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())

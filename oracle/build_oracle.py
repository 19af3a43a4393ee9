"""Build the C oracle (test infrastructure) into oracle/build/libes_oracle.so with gcc.

The reference is Python only, so there is no reference build under oracle/_ref."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "build", "libes_oracle.so")


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-pthread", "-Wall",
           os.path.join(HERE, "es_oracle.c"), "-o", OUT + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())

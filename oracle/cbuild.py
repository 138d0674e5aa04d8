"""Build recipe of the C oracle (oracle/csrc/oracle.c -> oracle/_ref/liboracle.so).

Test infrastructure only.  The reference (soundersim) is pure Python/JAX:
it has no C/C++ sources to compile, so there is no reference build; the C
file here is this project's own float64 restatement, used as the timed CPU
baseline ("port") and cross-checked against the NumPy oracle.
"""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "oracle.c")
OUT = os.path.join(HERE, "_ref", "liboracle.so")


def build(reference=False):
    """Compile the C oracle (gcc -O3 -fopenmp -ffp-contract=off)."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    subprocess.run(["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                    "-o", tmp, SRC, "-lm"], check=True)
    os.replace(tmp, OUT)
    return OUT

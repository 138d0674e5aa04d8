"""Copies the reference's simc golden fixtures (``tests/fixtures/{flat,
tilted,hill,sinusoid,crater}.{npz,json}`` of soundersim) into
tests/golden/simc/ unchanged, byte for byte, and records their SHA-256 in
tests/golden/simc/SHA256SUMS.  They are data -- simc's cluttergrams (4330
samples x 20 traces, float32), its first-return bins/positions and the
nav track in ECEF -- not source.  Run from the repository root:

    python tests/golden/make_simc_fixtures.py /root/reference/tests/fixtures
"""

import hashlib
import os
import shutil
import sys

NAMES = ("flat", "tilted", "hill", "sinusoid", "crater")


def main(src):
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "simc")
    os.makedirs(dst, exist_ok=True)
    lines = []
    for n in NAMES:
        for ext in ("npz", "json"):
            f = os.path.join(src, f"{n}.{ext}")
            shutil.copyfile(f, os.path.join(dst, f"{n}.{ext}"))
            with open(f, "rb") as fh:
                lines.append(f"{hashlib.sha256(fh.read()).hexdigest()}  {n}.{ext}")
    with open(os.path.join(dst, "SHA256SUMS"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv[1])

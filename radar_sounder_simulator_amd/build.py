"""In-tree build of the gfx950 HIP library (``libsoundersim_hip.so``).

``hipcc`` cross-compiles for gfx950 without a GPU.  Flags:
  -O3 -std=c++17 -fPIC -shared
  -munsafe-fp-atomics          hardware float atomics (ds_add_f64/f32,
                               global_atomic_add_f64) instead of CAS loops;
                               the outputs are ordinary device allocations
  -Xclang -target-feature -Xclang -packed-fp32-ops
                               no packed FP32 (v_pk_*_f32) instructions
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsoundersim_hip.so")
SOURCES = ["ss_kernels.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops"]


def build_library(verbose=False):
    """Compile every HIP source into the in-tree shared library."""
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libsoundersim_hip.so")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))

"""CPU SIMT build of the gfx950 kernels (tests/native/simt/hip/hip_runtime.h)
-- test infrastructure only.  ``lib(asan=False)`` compiles
radar_sounder_simulator_amd/csrc/ss_kernels.hip with g++ against the
emulation header into tests/native/_build/ and returns a ctypes handle with
the C-ABI signatures declared; "device" pointers are host pointers (CPU
torch tensors)."""

import ctypes
import os
import subprocess

from radar_sounder_simulator_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "radar_sounder_simulator_amd", "csrc")
HDR = os.path.join(HERE, "simt", "hip", "hip_runtime.h")


def build(asan=False):
    out = os.path.join(HERE, "_build", "libss_sim_asan.so" if asan else "libss_sim.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    deps = [HDR, os.path.join(ROOT, "include", "soundersim_hip.h")] + [
        os.path.join(CSRC, f) for f in ("ss_kernels.hip", "ss_math.h", "ss_refract.h",
                                        "ss_joint.h")]
    if os.path.exists(out) and all(os.path.getmtime(out) >= os.path.getmtime(d) for d in deps):
        return out
    tmp = f"{out}.{os.getpid()}.tmp"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-D__HIPCC__",
           "-I", os.path.join(HERE, "simt"), "-x", "c++", os.path.join(CSRC, "ss_kernels.hip"),
           "-o", tmp, "-lpthread"]
    if asan:
        cmd[1:1] = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)  # atomic: parallel test workers never load a partial file
    return out


_h = {}


def lib(asan=False):
    """Configured handle; use it with ``_lib.use(handle)`` to route the
    package's device layer to it (host tensors, device="cpu")."""
    if asan not in _h:
        h = _lib.configure(ctypes.CDLL(build(asan)))
        h.host_memory = True
        _h[asan] = h
    return _h[asan]

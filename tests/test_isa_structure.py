"""Compiled-code guard for the persistent kernels (CPU only: hipcc builds
the gfx950 device assembly without a GPU).

In a persistent kernel one lane takes the next work id from a global counter
(``if (tid == 0) s_w = atomicAdd(counter, 1)``), then the workgroup reads it
back between two barriers.  If the compiler lets a back edge re-enter that
barrier/read sequence without passing the atomic -- it did once, merging a
``tid == 0`` store at the end of an iteration with the ``tid == 0`` counter
update -- the lanes re-read a stale id and repeat one work item forever on
the device, while a lane-level emulation of the source still runs.  This
test checks, in the assembly of every kernel with a counter atomic, that
the two barriers after the atomic are reached only through the loop header
that performs it."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "radar_sounder_simulator_amd", "csrc", "ss_kernels.hip")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("isa") / "k.s")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17",
                    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o", out],
                   check=True, capture_output=True)
    return open(out).read()


def _body(asm, name):
    i = asm.index(name + ":")
    return asm[i:asm.index(".Lfunc_end", i)].split("\n")


def test_work_counter_loops_reenter_through_the_atomic(asm):
    kernels = [k for k in re.findall(r"\.amdhsa_kernel (\S+)", asm)
               if "global_atomic_add_x2" in "\n".join(_body(asm, k))]
    assert any("batch_kernel" in k for k in kernels)
    for k in kernels:
        lines = _body(asm, k)
        atom = next(i for i, l in enumerate(lines) if "global_atomic_add_x2" in l)
        bars = [i for i, l in enumerate(lines) if "s_barrier" in l and i > atom][:2]
        assert len(bars) == 2, k
        hdr = max(i for i, l in enumerate(lines) if "Loop Header" in l and i < atom)
        # no loop may start between the atomic's loop header and the barriers
        assert not any("Loop Header" in l for l in lines[hdr + 1:bars[1]]), k
        labels = {m.group(1): i for i, l in enumerate(lines)
                  if (m := re.match(r"^(\.LBB\d+_\d+):", l))}
        head = [lab for lab, i in labels.items() if hdr - 2 <= i <= hdr]
        inside = {lab for lab, i in labels.items() if hdr - 2 <= i <= bars[1]} - set(head)
        for i, l in enumerate(lines):
            m = re.search(r"s_(?:cbranch_\w+|branch)\s+(\.LBB\d+_\d+)", l)
            assert not (m and m.group(1) in inside and i > bars[1]), (k, l.strip())

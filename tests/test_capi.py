"""The C-ABI library: it loads without a GPU, exports every function that
include/soundersim_hip.h declares, and rejects bad arguments with SS_EINVAL /
SS_ERANGE before any device work (host-side validation only; no kernel is
launched here)."""

import ctypes
import os
import re

import numpy as np
import pytest

from radar_sounder_simulator_amd import _lib

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "soundersim_hip.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(ss_\w+)\s*\(", src, re.M)))


def test_header_declares_the_bound_symbols():
    assert declared() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.ss_version()


def test_struct_layouts_match_header():
    # doubles first then int32s: no padding; sizes pinned against the header
    assert ctypes.sizeof(_lib.SurfaceParams) == 13 * 8 + 6 * 4
    assert ctypes.sizeof(_lib.RefractedParams) == 14 * 8 + 12 * 4
    assert ctypes.sizeof(_lib.Job) == _lib.load().ss_sizeof_job() == 400


def test_invalid_arguments_rejected_on_host():
    lib = _lib.load()
    p = _lib.SurfaceParams()
    p.n_samples, p.dt, p.c = 64, 1e-8, 3e8
    args = [ctypes.byref(p), None, 0, 0, None, 0, 0, None, None, None, None, None, None]
    assert lib.ss_surface_field(7, *args) == -1
    assert b"precision" in lib.ss_last_error()
    p.n_samples = 0
    assert lib.ss_surface_field(0, *args) == -1
    p.n_samples = 64
    p.flags = _lib.SS_ROUGH
    assert lib.ss_surface_field(0, *args) == -1           # no constants table
    p.flags = 0
    p.pattern = 5
    assert lib.ss_surface_field(0, *args) == -1           # tabulated without table
    r = _lib.RefractedParams()
    r.n_samples, r.dt, r.c, r.n_crossed = 64, 1e-8, 3e8, 0
    rargs = [ctypes.byref(r), None, 0, 0, None, 0, 0, None, None, None, None, None, None,
             None, None, None, 0]
    assert lib.ss_refracted_field(0, *rargs, None) == -3          # n_crossed out of range
    r.n_crossed, r.flags = 1, _lib.SS_ROUGH                  # incoherent + roughness
    assert lib.ss_refracted_field(0, *rargs, None) == -1
    # mixed tier without its redo workspace: rejected before any device work
    r.flags = _lib.SS_COHERENT
    dummy = 4096                                            # never dereferenced
    wargs = [ctypes.byref(r), dummy, 10, 10, dummy, 4, 4, dummy, dummy, dummy, None, None,
             dummy, dummy, None]
    assert lib.ss_refracted_field(1, *wargs, None, 0, None) == -1
    assert b"workspace" in lib.ss_last_error()
    assert lib.ss_refracted_field(1, *wargs, dummy, 100, None) == -1
    assert lib.ss_refracted_field(1, *wargs, dummy + 4, 1 << 20, None) == -1   # alignment
    assert b"aligned" in lib.ss_last_error()
    assert lib.ss_range_compress(None, 1, 8, None, 1, None, None) == -1
    assert lib.ss_range_compress_fft(None, 1, 8, None, 1, 512, None, None) == -1
    assert lib.ss_range_compress_fft(4096, 1, 8, 4096, 1, 500, 8192, None) == -3   # nfft
    assert lib.ss_pulse_spectrum(4096, 300, 512, 4096, None) == -3                  # 601 taps
    assert lib.ss_compress_fft_size(4096, 960) == 4096      # 1921 taps: two 4096-point blocks
    assert lib.ss_compress_fft_size(4096, 5000) == 0        # too long: direct summation
    assert lib.ss_compress_fft_size(300, 20) == 256
    assert lib.ss_batch_field(0, None, 0, None, None, 0, 8, 0, None) == 0   # empty batch
    assert lib.ss_batch_field(0, None, 1, None, None, 0, 3, 0, None) == -1  # bad tile
    j = _lib.Job()
    j.kind = 7
    assert lib.ss_batch_field(0, ctypes.byref(j), 1, 4096, 4096, 1 << 20, 8, 0, None) == -1
    assert b"kind" in lib.ss_last_error()


def test_workspace_sizes():
    lib = _lib.load()
    # fp64 needs none; mixed: 256-byte header + one flag byte per (16-trace,
    # 256-facet) work item rounded up to 8 + 4 * 8 * 16 mask bytes per item
    # -- 1000 facets x 20 traces = 4 x 2 items
    assert lib.ss_refracted_workspace_bytes(0, 1000, 20) == 0
    assert lib.ss_refracted_workspace_bytes(1, 1000, 20) == 256 + 8 + 8 * 512
    assert lib.ss_refracted_workspace_bytes(1, 0, 20) == 0
    j = _lib.Job()
    j.kind, j.n_facets, j.n_traces = 1, 1000, 20
    assert lib.ss_batch_workspace_bytes(0, ctypes.byref(j), 1, 8) == 256
    assert lib.ss_batch_workspace_bytes(1, ctypes.byref(j), 1, 8) == 256 + 16 + 12 * 256
    assert lib.ss_batch_workspace_bytes(1, ctypes.byref(j), 1, 16) == 256 + 8 + 8 * 512
    assert lib.ss_batch_workspace_bytes(1, ctypes.byref(j), 1, 5) == -1


def test_fill_constants():
    c = _lib.constants_host()
    assert c.shape == (_lib.CONST_COUNT,) and np.isfinite(c).all()


def test_noalloc_rejects_calls_beyond_its_scratch():
    """ss_refracted_field_noalloc refuses, before any device work, a call
    whose redo record would exceed SS_NOALLOC_SCRATCH_BYTES."""
    lib = _lib.load()
    r = _lib.RefractedParams()
    r.n_samples, r.dt, r.c, r.n_crossed = 64, 1e-8, 3e8, 1
    n = 1 << 22       # 16384 facet tiles x 524288 trace tiles: far beyond 32 MiB
    assert lib.ss_refracted_workspace_bytes(1, n, n) > _lib.NOALLOC_SCRATCH_BYTES
    rc = lib.ss_refracted_field_noalloc(1, ctypes.byref(r), None, n, n, None, n, n, None, None,
                                        None, None, None, None, None, None, None)
    assert rc == -3
    assert b"SS_NOALLOC_SCRATCH_BYTES" in lib.ss_last_error()


def _job(kind, flags=0, refraction=0, n_crossed=1, pattern=0):
    """A host job record with placeholder (never dereferenced) pointers."""
    j = _lib.Job()
    j.kind = kind
    j.facet_stride = j.n_facets = 1000
    j.trace_stride = j.n_traces = 20
    p = ctypes.c_void_p(4096)
    j.facets = j.traces = j.field = j.dropped = j.constants = p
    j.iface_desc = j.tables = j.legs = j.pattern_table = p
    prm = j.surface if kind == 0 else j.refracted
    prm.n_samples, prm.dt, prm.c, prm.flags = 256, 1e-8, 3e8, flags
    prm.pattern, prm.pat_n = pattern, (181 if pattern else 0)
    if kind == 1:
        prm.refraction, prm.n_crossed = refraction, n_crossed
        prm.joint_newton, prm.joint_backtrack = 6, 4
    return j


def _pack(*jobs):
    lib = _lib.load()
    arr = (_lib.Job * len(jobs))(*jobs)
    buf = (ctypes.c_uint8 * lib.ss_batch_table_bytes(len(jobs)))()
    items = (ctypes.c_int64 * 2)()
    cls = ctypes.c_int()
    rc = lib.ss_batch_pack(arr, len(jobs), 16, buf, len(buf), items, ctypes.byref(cls))
    assert rc == 0, lib.ss_last_error()
    return cls.value, (items[0], items[1])


def test_batch_feature_class_rules():
    """Feature class 0 (compile-time flags in the mixed launches) takes only
    smooth, isotropic, plain jobs: coherent refracted jobs without split
    sides, diffuse term, per-facet gamma or forced redo, one crossing per
    sequential job and joint stacks of at most three; anything else of
    that kind is class 1; roughness beyond the area-only D_Phi or a
    non-tabulated pattern is class 2."""
    coh = _lib.SS_COHERENT
    assert _pack(_job(0)) == (0, (8, 0))          # 4 facet tiles x 2 trace tiles
    assert _pack(_job(0, _lib.SS_TAPER))[0] == 0
    assert _pack(_job(0, _lib.SS_SPLIT_SIDES))[0] == 1
    assert _pack(_job(0, _lib.SS_GAMMA_FACET))[0] == 1
    assert _pack(_job(1, coh)) == (0, (0, 8))
    assert _pack(_job(1, coh | _lib.SS_TAPER | _lib.SS_CROSS_ROUGH))[0] == 0
    assert _pack(_job(1, 0))[0] == 1               # incoherent refracted
    assert _pack(_job(1, coh | _lib.SS_DIFFUSE))[0] == 1
    assert _pack(_job(1, coh | _lib.SS_REDO_ALL))[0] == 1
    assert _pack(_job(1, coh, refraction=0, n_crossed=2))[0] == 1
    assert _pack(_job(1, coh, refraction=1, n_crossed=2))[0] == 0
    assert _pack(_job(1, coh, refraction=1, n_crossed=3))[0] == 0
    assert _pack(_job(1, coh, refraction=1, n_crossed=4))[0] == 1
    assert _pack(_job(1, coh, refraction=2, n_crossed=8))[0] == 0
    assert _pack(_job(0, pattern=5))[0] == 1        # tabulated pattern
    assert _pack(_job(0, pattern=1))[0] == 2        # dipole
    # the class of a table is its highest job's
    assert _pack(_job(0), _job(1, coh, refraction=1, n_crossed=4))[0] == 1
    assert _pack(_job(0), _job(1, coh), _job(0, pattern=1))[0] == 2

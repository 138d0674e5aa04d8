"""ctypes binding of the C ABI declared in include/soundersim_hip.h.

The shared library ``libsoundersim_hip.so`` is built in-tree by
``radar_sounder_simulator_amd.build.build_library()``.  There is no CPU
fallback: if the library is missing, every entry point raises.
"""

import contextlib
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsoundersim_hip.so")

SS_OK = 0
SS_PREC_FP64 = 0
SS_PREC_MIXED = 1
SS_SPLIT_SIDES = 0x001
SS_INTERP_BINS = 0x002
SS_TAPER = 0x004
SS_D_PHI_AREA = 0x008
SS_ROUGH = 0x010
SS_GAMMA_FACET = 0x020
SS_COHERENT = 0x040
SS_CROSS_ROUGH = 0x080
SS_DIFFUSE = 0x100
SS_REDO_ALL = 0x200

PATTERNS = {"isotropic": 0, "dipole": 1, "array": 2, "array_tapered": 3,
            "finite_dipole": 4, "tabulated": 5}
ACFS = {"gaussian": 0, "exponential": 1}

FROW = dict(CX=0, NX=3, AREA=6, E1X=7, E2X=10, GAMMA=13, PH_RE=14, PH_IM=15,
            DIF_AMP=16, DIF_RE=17, DIF_IM=18, COUNT=19)
TROW = dict(PX=0, UX=3, RREF=6, PV0=7, PV1=10, COUNT=13)
IDESC = dict(NY=0, NX=1, COEF=2, MP=8, MN=11, SIGC=14, TABOFF=15, COUNT=16)
BIN_INVALID = -2147483648
MAX_CROSSED = 64
MAX_JOINT = 32
BATCH_TABLE_HEADER = 256
MAX_TERMS = 300
CONST_COUNT = 33 + MAX_TERMS + 1

EXPORTED = ("ss_version", "ss_last_error", "ss_device_count",
            "ss_fill_constants", "ss_surface_field", "ss_refracted_field",
            "ss_refracted_workspace_bytes", "ss_batch_field", "ss_batch_workspace_bytes",
            "ss_sizeof_job", "ss_range_compress", "ss_compress_fft_size", "ss_pulse_spectrum",
            "ss_range_compress_fft", "ss_batch_table_bytes", "ss_batch_pack",
            "ss_batch_step_workspace_bytes", "ss_batch_step", "ss_facets_from_heights",
            "ss_facets_workspace_bytes", "ss_trace_diagnostics", "ss_incoherent_power",
            "ss_slow_time_boxcar", "ss_focused_sar", "ss_refracted_field_noalloc",
            "ss_batch_overflowed", "ss_gaussian_heights",
            "ss_gaussian_heights_workspace_bytes", "ss_keyed_phasors")
NOALLOC_SCRATCH_BYTES = 32 << 20


class SurfaceParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("gamma", ctypes.c_double),
                ("t0", ctypes.c_double), ("dt", ctypes.c_double),
                ("c", ctypes.c_double), ("taper_s", ctypes.c_double),
                ("sigma", ctypes.c_double), ("corr_len", ctypes.c_double),
                ("pat_a", ctypes.c_double), ("pat_b", ctypes.c_double * 4),
                ("n_samples", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("n_terms", ctypes.c_int32), ("acf", ctypes.c_int32),
                ("pattern", ctypes.c_int32), ("pat_n", ctypes.c_int32)]


class RefractedParams(ctypes.Structure):
    _fields_ = [("k0", ctypes.c_double), ("gamma", ctypes.c_double),
                ("t0", ctypes.c_double), ("dt", ctypes.c_double),
                ("c", ctypes.c_double), ("taper_s", ctypes.c_double),
                ("sigma", ctypes.c_double), ("corr_len", ctypes.c_double),
                ("n_exp", ctypes.c_double), ("pat_a", ctypes.c_double),
                ("pat_b", ctypes.c_double * 4),
                ("n_samples", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("n_crossed", ctypes.c_int32), ("n_terms", ctypes.c_int32),
                ("acf", ctypes.c_int32), ("pattern", ctypes.c_int32),
                ("pat_n", ctypes.c_int32), ("refraction", ctypes.c_int32),
                ("k_pad", ctypes.c_int32), ("joint_newton", ctypes.c_int32),
                ("joint_backtrack", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class IncoherentParams(ctypes.Structure):
    _fields_ = [("t0", ctypes.c_double), ("dt", ctypes.c_double),
                ("c", ctypes.c_double), ("pat_a", ctypes.c_double),
                ("pat_b", ctypes.c_double * 4),
                ("n_samples", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("pattern", ctypes.c_int32), ("pat_n", ctypes.c_int32)]


class Job(ctypes.Structure):
    """Mirror of ``ss_job`` (one layer of one request in a batched launch)."""
    _fields_ = [("kind", ctypes.c_int32), ("precision_unused", ctypes.c_int32),
                ("facet_stride", ctypes.c_int64), ("n_facets", ctypes.c_int64),
                ("trace_stride", ctypes.c_int64), ("n_traces", ctypes.c_int64),
                ("facets", ctypes.c_void_p), ("traces", ctypes.c_void_p),
                ("field", ctypes.c_void_p), ("dropped", ctypes.c_void_p),
                ("iface_desc", ctypes.c_void_p), ("tables", ctypes.c_void_p),
                ("legs", ctypes.c_void_p), ("pattern_table", ctypes.c_void_p),
                ("constants", ctypes.c_void_p), ("surface", SurfaceParams),
                ("refracted", RefractedParams)]


_lib = None
_override = None


@contextlib.contextmanager
def use(lib):
    """Route every call of the package to ``lib`` (a handle prepared by
    configure(), e.g. a CPU build of the same kernels) inside the block."""
    global _override
    prev, _override = _override, lib
    try:
        yield lib
    finally:
        _override = prev


def load():
    """Load (once) and return the ctypes library; raises if it is missing."""
    global _lib
    if _override is not None:
        return _override
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP library not built: {LIB_PATH} is missing (run "
            "radar_sounder_simulator_amd.build.build_library())")
    _lib = configure(ctypes.CDLL(LIB_PATH))
    return _lib


def configure(lib):
    """Declare the C-ABI signatures on a loaded library handle."""
    vp, i64, i32, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
    lib.ss_version.restype = ctypes.c_char_p
    lib.ss_last_error.restype = ctypes.c_char_p
    lib.ss_device_count.restype = ctypes.c_int
    lib.ss_fill_constants.argtypes = [dp]
    lib.ss_fill_constants.restype = ctypes.c_int
    lib.ss_surface_field.argtypes = [i32, ctypes.POINTER(SurfaceParams), dp, i64,
                                     i64, dp, i64, i64, dp, dp, dp, dp, dp, vp]
    lib.ss_surface_field.restype = ctypes.c_int
    lib.ss_incoherent_power.argtypes = [i32, ctypes.POINTER(IncoherentParams), dp, i64,
                                        i64, dp, i64, i64, dp, dp, dp, dp, vp]
    lib.ss_incoherent_power.restype = ctypes.c_int
    lib.ss_refracted_field.argtypes = [i32, ctypes.POINTER(RefractedParams), dp,
                                       i64, i64, dp, i64, i64, dp, dp, dp, dp,
                                       dp, dp, dp, dp, dp, i64, vp]
    lib.ss_refracted_field.restype = ctypes.c_int
    lib.ss_refracted_field_noalloc.argtypes = [i32, ctypes.POINTER(RefractedParams), dp,
                                               i64, i64, dp, i64, i64, dp, dp, dp, dp,
                                               dp, dp, dp, dp, vp]
    lib.ss_refracted_field_noalloc.restype = ctypes.c_int
    lib.ss_batch_overflowed.argtypes = [dp, ctypes.POINTER(ctypes.c_int), vp]
    lib.ss_batch_overflowed.restype = ctypes.c_int
    lib.ss_refracted_workspace_bytes.argtypes = [i32, i64, i64]
    lib.ss_refracted_workspace_bytes.restype = ctypes.c_int64
    lib.ss_batch_field.argtypes = [i32, ctypes.POINTER(Job), i32, dp, dp, i64, i32, i32, vp]
    lib.ss_batch_field.restype = ctypes.c_int
    lib.ss_batch_workspace_bytes.argtypes = [i32, ctypes.POINTER(Job), i32, i32]
    lib.ss_batch_workspace_bytes.restype = ctypes.c_int64
    lib.ss_sizeof_job.restype = ctypes.c_int64
    lib.ss_range_compress.argtypes = [dp, i64, i64, dp, i64, dp, vp]
    lib.ss_range_compress.restype = ctypes.c_int
    lib.ss_compress_fft_size.argtypes = [i64, i64]
    lib.ss_compress_fft_size.restype = ctypes.c_int64
    lib.ss_pulse_spectrum.argtypes = [dp, i64, i64, dp, vp]
    lib.ss_pulse_spectrum.restype = ctypes.c_int
    lib.ss_range_compress_fft.argtypes = [dp, i64, i64, dp, i64, i64, dp, vp]
    lib.ss_range_compress_fft.restype = ctypes.c_int
    lib.ss_batch_table_bytes.argtypes = [i32]
    lib.ss_batch_table_bytes.restype = ctypes.c_int64
    lib.ss_batch_pack.argtypes = [ctypes.POINTER(Job), i32, i32, dp, i64,
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
    lib.ss_batch_pack.restype = ctypes.c_int
    lib.ss_batch_step_workspace_bytes.argtypes = [i32, i64, i32]
    lib.ss_batch_step_workspace_bytes.restype = ctypes.c_int64
    lib.ss_batch_step.argtypes = [i32, dp, dp, i64, i64, i32, i32, i32, vp]
    lib.ss_batch_step.restype = ctypes.c_int
    lib.ss_facets_from_heights.argtypes = [dp, i64, i64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, dp, i64, dp, dp, dp, i64, vp]
    lib.ss_facets_from_heights.restype = ctypes.c_int
    lib.ss_facets_workspace_bytes.argtypes = [i64, i64]
    lib.ss_facets_workspace_bytes.restype = ctypes.c_int64
    d = ctypes.c_double
    lib.ss_trace_diagnostics.argtypes = [dp, i64, i64, dp, dp, i64, i64, d, d, i64, d,
                                         dp, dp, dp, dp, dp, vp]
    lib.ss_trace_diagnostics.restype = ctypes.c_int
    lib.ss_slow_time_boxcar.argtypes = [dp, i64, i64, i64, i64, i64, d, dp, vp]
    lib.ss_slow_time_boxcar.restype = ctypes.c_int
    lib.ss_focused_sar.argtypes = [i32, dp, i64, i64, dp, dp, d, d, d, i32, dp, vp]
    lib.ss_focused_sar.restype = ctypes.c_int
    u64 = ctypes.c_uint64
    lib.ss_gaussian_heights.argtypes = [u64, i64, i64, d, d, d, d, dp, dp, i64, vp]
    lib.ss_gaussian_heights.restype = ctypes.c_int
    lib.ss_gaussian_heights_workspace_bytes.argtypes = [i64, i64]
    lib.ss_gaussian_heights_workspace_bytes.restype = ctypes.c_int64
    lib.ss_keyed_phasors.argtypes = [u64, i64, i64, i64, i64, dp, i64, vp]
    lib.ss_keyed_phasors.restype = ctypes.c_int
    return lib


def check(rc, what):
    if rc != SS_OK:
        msg = load().ss_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def constants_host():
    """The kernel constants table (Weideman coefficients, lgamma) as numpy."""
    import numpy as np
    out = np.zeros(CONST_COUNT, np.float64)
    check(load().ss_fill_constants(out.ctypes.data), "ss_fill_constants")
    return out

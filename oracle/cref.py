"""ctypes wrapper of the C oracle (oracle/csrc/oracle.c) -- test / CPU-baseline
infrastructure only.  Same arguments and outputs as the NumPy oracle's
``coherent_cluttergram`` / ``refracted_cluttergram`` for the feature subset
the C file implements (area-only roughness, taper, per-facet gamma,
isotropic / dipole / tabulated patterns)."""

import ctypes
import os

import numpy as np

from . import cbuild
from .coherent import reference_range
from .multilayer import grid_consts


class _P(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("gamma", ctypes.c_double),
                ("t0", ctypes.c_double), ("dt", ctypes.c_double),
                ("c", ctypes.c_double), ("taper_s", ctypes.c_double),
                ("sigma", ctypes.c_double), ("lc", ctypes.c_double),
                ("n", ctypes.c_long), ("n_terms", ctypes.c_int),
                ("acf", ctypes.c_int), ("pat", ctypes.c_int),
                ("ntab", ctypes.c_int), ("nthreads", ctypes.c_int),
                ("tab", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(cbuild.OUT):
            cbuild.build()
        _lib = ctypes.CDLL(cbuild.OUT)
    return _lib


def _a(x, dt=np.float64):
    return np.ascontiguousarray(np.asarray(x, dt))


def _ptr(a):
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def _params(k, gamma, t0, dt, n, c, taper_s, roughness, d_phi_area, pattern, nthreads):
    p = _P()
    p.k, p.gamma, p.t0, p.dt, p.c = k, float(np.ndim(gamma) == 0 and gamma or 0.0), t0, dt, c
    p.n = n
    p.taper_s = -1.0 if taper_s is None else taper_s
    p.acf = -1
    ph = None
    if roughness is not None:
        sigma, lc, phasors, n_terms, *rest = roughness
        acf = rest[0] if rest else "gaussian"
        if not d_phi_area:
            raise ValueError("the C oracle implements the area-only D_Phi only")
        p.sigma, p.lc, p.n_terms = sigma, lc, n_terms
        p.acf = 1 if acf == "exponential" else 0
        phv = np.asarray(phasors, np.complex128)
        ph = _a(np.stack([phv.real, phv.imag], -1))
    tab = pv = None
    p.pat = 0
    if pattern is not None and pattern[0] != "isotropic":
        kind, pvv, pa, pb = pattern
        if kind == "dipole":
            p.pat = 1
        elif kind == "tabulated":
            p.pat = 5
            tab = _a(np.concatenate([np.asarray(pa, np.float64), np.asarray(pb, np.float64)]))
            p.ntab = len(pa)
            p.tab = tab.ctypes.data
        else:
            raise ValueError(f"C oracle: pattern {kind!r} not implemented")
        pv = _a(pvv)
    p.nthreads = nthreads
    return p, ph, tab, pv


def coherent_cluttergram(positions, u_ct, centers, normals, areas, e1, e2, *, k, gamma,
                         t0, dt, n_samples, c, pattern=None, roughness=None, taper_s=None,
                         d_phi_area=False, nthreads=1, return_bins=False):
    """Surface field (T, n) complex128 and dropped (T,); ``return_bins``
    adds the (T, F) int32 floor bins of every pair."""
    pos = _a(positions)
    T, F = len(pos), len(centers)
    p, ph, tab, pv = _params(k, gamma, t0, dt, n_samples, c, taper_s, roughness,
                             d_phi_area, pattern, nthreads)
    gf = _a(gamma) if np.ndim(gamma) else None
    rref = _a(reference_range(pos, centers))
    field = np.zeros((T, n_samples, 2))
    drop = np.zeros(T)
    C, N, A, E1, E2 = (_a(v) for v in (centers, normals, areas, e1, e2))
    bins = np.empty((T, F), np.int32) if return_bins else None
    lib().oracle_surface(ctypes.byref(p), _ptr(pos), _ptr(pv), _ptr(rref), ctypes.c_long(T),
                         _ptr(C), _ptr(N), _ptr(A), _ptr(E1), _ptr(E2), _ptr(gf), _ptr(ph),
                         ctypes.c_long(F), _ptr(field), _ptr(drop), _ptr(bins))
    out = field[..., 0] + 1j * field[..., 1]
    return (out, drop, bins) if return_bins else (out, drop)


def refracted_cluttergram(positions, u_ct, target, crossed, eps_leg, att_leg, *, t0, dt,
                          n_samples, c, gamma=0.0, k0=None, pattern=None, roughness=None,
                          taper_s=None, d_phi_area=False, nthreads=1, mode="coherent",
                          return_bins=False):
    """Sequential-chain refracted field (T, n) complex128 and dropped (T,);
    ``return_bins`` adds the (T, F) int32 floor bins, INT32_MIN where the
    path is invalid."""
    if mode != "coherent":
        raise ValueError("C oracle: coherent mode only")
    pos = _a(positions)
    T, F = len(pos), len(target.centers)
    p, ph, tab, pv = _params(k0, gamma, t0, dt, n_samples, c, taper_s, roughness,
                             d_phi_area, pattern, nthreads)
    gf = _a(gamma) if np.ndim(gamma) else None
    desc = np.zeros((len(crossed), 16))
    tabs = []
    off = 0
    for i, f in enumerate(crossed):
        tb, coef, mp, mn = grid_consts(f)
        ny, nx = f.grid_shape
        desc[i, 0], desc[i, 1] = ny, nx
        desc[i, 2:8] = coef.ravel()
        desc[i, 8:11], desc[i, 11:14] = mp, mn
        desc[i, 15] = off
        off += ny * nx
        tabs.append(tb.reshape(-1, 6))
    tables = _a(np.concatenate(tabs))
    eps = _a(eps_leg)
    nl = _a(np.sqrt(eps))
    att = _a(att_leg)
    field = np.zeros((T, n_samples, 2))
    drop = np.zeros(T)
    C, N, A, E1, E2 = (_a(v) for v in (target.centers, target.normals, target.areas,
                                       target.e1, target.e2))
    desc = _a(desc)
    bins = np.empty((T, F), np.int32) if return_bins else None
    lib().oracle_refracted(ctypes.byref(p), _ptr(pos), _ptr(pv), ctypes.c_long(T), _ptr(C),
                           _ptr(N), _ptr(A), _ptr(E1), _ptr(E2), _ptr(gf), _ptr(ph),
                           ctypes.c_long(F), ctypes.c_int(len(crossed)), _ptr(desc),
                           _ptr(tables), _ptr(nl), _ptr(eps), _ptr(att), _ptr(field),
                           _ptr(drop), _ptr(bins))
    out = field[..., 0] + 1j * field[..., 1]
    return (out, drop, bins) if return_bins else (out, drop)

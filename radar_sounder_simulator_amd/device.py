"""Device-resident scenes and kernel launches (torch tensors as HBM buffers).

A ``DeviceFacets`` holds one interface's facets in HBM as the float64
structure-of-arrays the kernels read (rows of include/soundersim_hip.h
SS_FROW_*), in Morton (Z-curve) order of the horizontal centre coordinates,
so that 256 consecutive facets -- one workgroup tile -- form a compact 2-D
patch whose pairs fall in a narrow band of fast-time bins for any trace.
The facet sum is order independent up to rounding (the reference itself
reorders facets along track, ``kernels/geometry.py:48-60``).

A ``DeviceTraces`` holds the per-trace records (positions, cross-track
vectors, surface reference range, antenna pattern vectors).

``surface_field`` / ``refracted_field`` / ``range_compress`` enqueue the
kernels on a stream and return device tensors; nothing is copied to the
host.  These are the allocation-aware building blocks under the reference-
signature shims in ``kernels.py``, the batch runner and the serving loop.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import FROW, IDESC, TROW
from .roughness import KeyedPhasors, key_base
from .scene import SeededLattice


def _dev(device):
    """torch device for HBM buffers ("cpu" is accepted for host-side
    inspection of the buffer layout; the kernels need a GPU tensor)."""
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def _stream_handle(stream, dev=None):
    """HIP stream handle for a launch on tensors of ``dev``.  Host tensors
    are accepted only by a library build that runs on host memory (the CPU
    emulation the tests use); the HIP library refuses them here, before any
    host pointer could reach a kernel."""
    if dev is not None and torch.device(dev).type != "cuda":
        if not getattr(_lib.load(), "host_memory", False):
            raise RuntimeError(f"the HIP kernels need GPU tensors, got {dev} tensors")
        return ctypes.c_void_p(0)
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream.cuda_stream)


def morton_order(xy):
    """Stable argsort of 2-D points along a Z-order curve (16 bits/axis)."""
    xy = np.asarray(xy, np.float64)
    if len(xy) == 0:
        return np.zeros(0, np.int64)
    lo = xy.min(0)
    span = np.maximum(xy.max(0) - lo, 1e-12)
    q = np.minimum(((xy - lo) / span * 65535.0).astype(np.uint64), 65535)

    def spread(v):
        v = v & np.uint64(0xFFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1))
    return np.argsort(code, kind="stable")


_CONST_CACHE = {}


def constants(device=None):
    """Kernel constants table on ``device`` (cached)."""
    dev = _dev(device)
    key = str(dev)
    if key not in _CONST_CACHE:
        _CONST_CACHE[key] = torch.as_tensor(_lib.constants_host(), device=dev)
    return _CONST_CACHE[key]


class DeviceFacets:
    """One interface's facets resident in HBM (float64 SoA, Morton order)."""

    def __init__(self, centers, normals, areas, e1, e2, *, gamma=None,
                 phasors=None, diffuse_amp=None, diffuse_phasors=None,
                 device=None, order=None):
        centers = np.asarray(centers, np.float64)
        n = centers.shape[0]
        self.n = n
        self.order = morton_order(centers[:, :2]) if order is None else np.asarray(order)
        o = self.order
        buf = np.zeros((FROW["COUNT"], max(n, 1)), np.float64)
        buf[0:3, :n] = centers[o].T
        buf[3:6, :n] = np.asarray(normals, np.float64)[o].T
        buf[6, :n] = np.asarray(areas, np.float64)[o]
        buf[7:10, :n] = np.asarray(e1, np.float64)[o].T
        buf[10:13, :n] = np.asarray(e2, np.float64)[o].T
        self.has_gamma = gamma is not None
        if gamma is not None:
            buf[13, :n] = np.asarray(gamma, np.float64)[o]
        self.has_phasors = phasors is not None
        if phasors is not None:
            ph = np.asarray(phasors)[o]
            buf[14, :n] = ph.real
            buf[15, :n] = ph.imag
        self.has_diffuse = diffuse_amp is not None
        if diffuse_amp is not None:
            buf[16, :n] = np.asarray(diffuse_amp, np.float64)[o]
            dp = np.asarray(diffuse_phasors)[o]
            buf[17, :n] = dp.real
            buf[18, :n] = dp.imag
        self.device = _dev(device)
        self.buf = torch.as_tensor(buf, device=self.device)
        self.stride = buf.shape[1]

    @classmethod
    def from_facets(cls, f, **kw):
        return cls(f.centers, f.normals, f.areas, f.e1, f.e2, **kw)

    def nbytes(self):
        return self.buf.numel() * 8

    def unpermute_pairs(self, pairs):
        """(T, F) per-pair array in device order -> caller's facet order."""
        out = np.empty_like(pairs)
        out[:, self.order] = pairs
        return out


def tile_order(ny, nx):
    """Cell (grid-order index) at each facet position of the device-built
    order: 16x16-cell tiles, tiles row-major, cells row-major in a tile."""
    r, c = np.divmod(np.arange(ny * nx, dtype=np.int64), nx)
    tr, tc = r // 16, c // 16
    h = np.minimum(16, ny - tr * 16)
    w = np.minimum(16, nx - tc * 16)
    pos = tr * 16 * nx + tc * 16 * h + (r % 16) * w + (c % 16)
    order = np.empty(ny * nx, np.int64)
    order[pos] = np.arange(ny * nx, dtype=np.int64)
    return order


class BuiltLattice:
    """A lattice tessellated on the device: its facets (``fac``), the
    (sum A, sum A c, sum A n, sum c) statistics and, when the lattice is a
    crossed interface, its lookup-table rows (``table_off`` into the shared
    table; one row for a horizontal plane)."""

    def __init__(self, lat, fac, stats, table_off, horizontal):
        self.lattice, self.fac, self.stats = lat, fac, stats
        self.table_off, self.horizontal = table_off, horizontal
        ny, nx = lat.grid_shape
        self.ny, self.nx = ny, nx
        self.center_mean = stats[7:10] / (ny * nx)

    def z_max(self):
        """The highest facet centre (the reference places joint pads above
        it), reduced on the device; a lattice with host heights gives the
        host value."""
        if not isinstance(self.lattice, SeededLattice) or self.lattice._heights is not None:
            h = self.lattice.heights
            return float(((((h[:-1, :-1] + h[:-1, 1:]) + h[1:, :-1]) + h[1:, 1:]) / 4.0).max())
        return float(self.fac.buf[2, :self.fac.n].max())

    def desc_row(self):
        """Descriptor row (TABOFF/SIGC unset): grid shape, the exact affine
        (x, y, 1) -> (row, col) map of the regular grid (what the reference's
        least-squares fit recovers), the area-weighted mean plane."""
        lat = self.lattice
        d = np.zeros(IDESC["COUNT"], np.float64)
        A = self.stats[0]
        mp = self.stats[1:4] / A
        mn = self.stats[4:7] / np.linalg.norm(self.stats[4:7])
        if self.horizontal:
            d[IDESC["NY"]] = d[IDESC["NX"]] = 1
        else:
            s = float(lat.spacing)
            x0, y0 = map(float, lat.origin)
            hy, hx = self.ny / 2.0, self.nx / 2.0
            coef = np.array([[0.0, 1.0 / s], [1.0 / s, 0.0],
                             [hy - 0.5 - y0 / s, hx - 0.5 - x0 / s]])
            d[IDESC["NY"]], d[IDESC["NX"]] = self.ny, self.nx
            d[IDESC["COEF"]:IDESC["COEF"] + 6] = coef.ravel()
        d[IDESC["MP"]:IDESC["MP"] + 3] = mp
        d[IDESC["MN"]:IDESC["MN"] + 3] = mn
        return d


def lattice_heights(lat, device=None, stream=None):
    """(ny+1, nx+1) float64 heights of ``lat`` in HBM: a ``SeededLattice``
    drawn on the device (``ss_gaussian_heights``; a horizontal one filled),
    any other lattice uploaded."""
    dev = _dev(device)
    if not isinstance(lat, SeededLattice):
        return torch.as_tensor(np.ascontiguousarray(lat.heights, np.float64), device=dev)
    rows, cols = lat.ny + 1, lat.nx + 1
    if lat.sigma == 0.0:
        h = torch.empty((rows, cols), dtype=torch.float64, device=dev)
        if stream is None:
            return h.fill_(lat.mean)
        with torch.cuda.stream(stream):
            return h.fill_(lat.mean)
    lib = _lib.load()
    h = torch.empty((rows, cols), dtype=torch.float64, device=dev)
    ws = torch.empty(int(lib.ss_gaussian_heights_workspace_bytes(rows, cols)),
                     dtype=torch.uint8, device=dev)
    rc = lib.ss_gaussian_heights(int(key_base(lat.seed)), rows, cols, lat.spacing,
                                 lat.corr_len, lat.sigma, lat.mean, _ptr(h), _ptr(ws),
                                 ws.numel(), _stream_handle(stream, dev))
    _lib.check(rc, "ss_gaussian_heights")
    if dev.type == "cuda":
        # the workspace is freed on return: keep the caching allocator from
        # handing it to another stream before the launch has read it
        ws.record_stream(stream if stream is not None else torch.cuda.current_stream(dev))
    return h


def build_lattices(lattices, crossed, device=None, gammas=None, phasors=None,
                   n_pad_rows=0, stream=None):
    """Tessellate every distinct lattice (by identity) once on the device
    (``ss_facets_from_heights``): facets for each, and for those in
    ``crossed`` (a set of ids) their lookup rows in one shared table with
    ``n_pad_rows`` spare rows at its end.  ``gammas``/``phasors``: optional
    {id: per-cell array in grid order}.  Returns ({id: BuiltLattice}, table
    tensor)."""
    dev = _dev(device)
    lib = _lib.load()
    uniq = {}
    for lat in lattices:
        uniq.setdefault(id(lat), lat)
    hz = {k: lat.horizontal() for k, lat in uniq.items()}
    rows, offs = 0, {}
    for k, lat in uniq.items():
        if k in crossed:
            offs[k] = rows
            rows += 1 if hz[k] else len(lat)
    table = torch.zeros((max(rows + n_pad_rows, 1), 6), dtype=torch.float64, device=dev)
    st = _stream_handle(stream, dev)
    gammas, phasors = gammas or {}, phasors or {}
    out, pending = {}, []
    for k, lat in uniq.items():
        ny, nx = lat.grid_shape
        n = ny * nx
        h = lattice_heights(lat, dev, stream)
        buf = torch.zeros((FROW["COUNT"], n), dtype=torch.float64, device=dev)
        stats = torch.zeros(16, dtype=torch.float64, device=dev)
        ws = torch.empty(max(int(lib.ss_facets_workspace_bytes(ny, nx)), 8),
                         dtype=torch.uint8, device=dev)
        tab = None
        if k in crossed and not hz[k]:
            tab = table[offs[k]:offs[k] + n]
        rc = lib.ss_facets_from_heights(_ptr(h), ny, nx, float(lat.spacing),
                                        float(lat.origin[0]), float(lat.origin[1]), _ptr(buf),
                                        n, _ptr(tab), _ptr(stats), _ptr(ws), ws.numel(), st)
        _lib.check(rc, "ss_facets_from_heights")
        order = tile_order(ny, nx)
        fac = object.__new__(DeviceFacets)
        fac.n, fac.order, fac.device, fac.buf, fac.stride = n, order, dev, buf, n
        fac.has_gamma = k in gammas
        fac.has_phasors = k in phasors
        fac.has_diffuse = False
        if fac.has_gamma:
            buf[13] = torch.as_tensor(np.asarray(gammas[k], np.float64)[order], device=dev)
        if fac.has_phasors:
            ph = phasors[k]
            if isinstance(ph, KeyedPhasors):
                if ph.grid_shape != (ny, nx):
                    raise ValueError("keyed phasors of another grid shape")
                rc = lib.ss_keyed_phasors(int(key_base(ph.seed)), ny, nx, ph.key_origin[0],
                                          ph.key_origin[1], _ptr(buf), n, st)
                _lib.check(rc, "ss_keyed_phasors")
            else:
                ph = np.asarray(ph)[order]
                buf[14] = torch.as_tensor(np.ascontiguousarray(ph.real), device=dev)
                buf[15] = torch.as_tensor(np.ascontiguousarray(ph.imag), device=dev)
        pending.append((k, lat, fac, stats, h, ws))
    for k, lat, fac, stats, h, ws in pending:
        s_host = stats[:10].cpu().numpy()        # waits for the builds (stream order)
        out[k] = BuiltLattice(lat, fac, s_host, offs.get(k), hz[k])
        out[k].heights = h
        if k in crossed and hz[k]:
            ny, nx = lat.grid_shape
            z = lat.level()
            x = (0.5 - nx / 2.0) * lat.spacing + lat.origin[0]
            y = (0.5 - ny / 2.0) * lat.spacing + lat.origin[1]
            table[offs[k]] = torch.tensor([x, y, z, 0.0, 0.0, 1.0], dtype=torch.float64)
    return out, table


class DeviceTraces:
    """Per-trace records resident in HBM (rows SS_TROW_*)."""

    def __init__(self, positions, u_ct, r_ref=None, pv=None, device=None):
        pos = np.asarray(positions, np.float64)
        n = pos.shape[0]
        self.n = n
        buf = np.zeros((TROW["COUNT"], max(n, 1)), np.float64)
        buf[0:3, :n] = pos.T
        buf[3:6, :n] = np.asarray(u_ct, np.float64).reshape(n, 3).T
        if r_ref is not None:
            buf[6, :n] = r_ref
        if pv is not None:
            pv = np.asarray(pv, np.float64)
            if pv.ndim == 3:          # arrays: (T, 2, 3)
                buf[7:10, :n] = pv[:, 0].T
                buf[10:13, :n] = pv[:, 1].T
            else:
                buf[7:10, :n] = pv.reshape(n, 3).T
        self.device = _dev(device)
        self.buf = torch.as_tensor(buf, device=self.device)
        self.stride = buf.shape[1]


def pattern_setup(pattern, device):
    """(kind id, n, a, b[4], table tensor or None) from pattern_args output."""
    if pattern is None or pattern[0] == "isotropic":
        return 0, 0, 0.0, (0.0, 0.0, 0.0, 0.0), None
    kind, pv, pa, pb = pattern
    kid = _lib.PATTERNS[kind]
    b = [0.0, 0.0, 0.0, 0.0]
    table = None
    n = 0
    a = 0.0
    if kind == "array":
        a = float(pa)
        b = [float(v) for v in np.asarray(pb, np.float64)[:4]]
    elif kind == "array_tapered":
        w = np.asarray(pa, np.float64)
        n = w.shape[1]
        table = torch.as_tensor(np.concatenate([w[0], w[1]]), device=device)
        b = [float(v) for v in np.asarray(pb, np.float64)[:4]]
    elif kind == "finite_dipole":
        a = float(pa)
    elif kind == "tabulated":
        xp = np.asarray(pa, np.float64)
        fp = np.asarray(pb, np.float64)
        n = len(xp)
        table = torch.as_tensor(np.concatenate([xp, fp]), device=device)
    return kid, n, a, tuple(b), table


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def surface_params(facets, *, k, gamma, t0, dt, n_samples, c, split_sides=False,
                   interp_bins=False, pattern=None, roughness=None, taper_s=None,
                   d_phi_area=False):
    """(SurfaceParams, pattern table tensor or None) for the surface kernel.

    ``roughness``: None or (sigma, corr_len, n_terms, acf); the phasors live
    in ``facets``.  ``pattern``: None or a ``pattern_args`` tuple whose pv
    rows are in the traces buffer."""
    flags = 0
    flags |= _lib.SS_SPLIT_SIDES if split_sides else 0
    flags |= _lib.SS_INTERP_BINS if interp_bins else 0
    flags |= _lib.SS_TAPER if taper_s is not None else 0
    flags |= _lib.SS_D_PHI_AREA if d_phi_area else 0
    flags |= _lib.SS_GAMMA_FACET if facets.has_gamma else 0
    p = _lib.SurfaceParams()
    p.k, p.gamma, p.t0, p.dt, p.c = float(k), float(gamma), float(t0), float(dt), float(c)
    p.taper_s = 0.0 if taper_s is None else float(taper_s)
    p.n_samples = int(n_samples)
    if roughness is not None:
        sigma, lc, n_terms, acf = roughness
        if not facets.has_phasors:
            raise ValueError("roughness needs facets uploaded with phasors")
        flags |= _lib.SS_ROUGH
        p.sigma, p.corr_len, p.n_terms = float(sigma), float(lc), int(n_terms)
        p.acf = _lib.ACFS[acf]
    kid, pn, pa, pb, table = pattern_setup(pattern, facets.device)
    p.pattern, p.pat_n, p.pat_a = kid, pn, pa
    for i in range(4):
        p.pat_b[i] = pb[i]
    p.flags = flags
    return p, table


def surface_field(facets, traces, *, k, gamma, t0, dt, n_samples, c,
                  precision="mixed", split_sides=False, interp_bins=False,
                  pattern=None, roughness=None, taper_s=None, d_phi_area=False,
                  out=None, dropped=None, pair_bins=False, stream=None):
    """Enqueue the surface kernel.  Returns (field (T, n_seg, 2) float64,
    dropped (T,) float64[, pair_bins (T, F) int32], keep-alive) device tensors."""
    dev = facets.device
    p, table = surface_params(facets, k=k, gamma=gamma, t0=t0, dt=dt, n_samples=n_samples,
                              c=c, split_sides=split_sides, interp_bins=interp_bins,
                              pattern=pattern, roughness=roughness, taper_s=taper_s,
                              d_phi_area=d_phi_area)
    n_seg = (2 if split_sides else 1) * int(n_samples)
    T = traces.n
    if out is None:
        out = torch.empty((T, n_seg, 2), dtype=torch.float64, device=dev)
    if dropped is None:
        dropped = torch.empty((T,), dtype=torch.float64, device=dev)
    pb_t = (torch.empty((T, facets.n), dtype=torch.int32, device=dev)
            if pair_bins else None)
    cst = constants(dev)
    prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
    rc = _lib.load().ss_surface_field(
        prec, ctypes.byref(p), _ptr(facets.buf), facets.stride, facets.n,
        _ptr(traces.buf), traces.stride, T, _ptr(table), _ptr(cst), _ptr(out),
        _ptr(dropped), _ptr(pb_t), _stream_handle(stream, dev))
    _lib.check(rc, "ss_surface_field")
    if pair_bins:
        return out, dropped, pb_t, table
    return out, dropped, table


def incoherent_power(facets, traces, *, t0, dt, n_samples, c, precision="mixed",
                     split_sides=False, pattern=None, out=None, dropped=None,
                     pair_bins=False, stream=None):
    """Enqueue the incoherent kernel.  Returns (power (T, n_seg) float64,
    dropped (T,) float64[, pair_bins (T, F) int32], keep-alive) device tensors."""
    dev = facets.device
    p = _lib.IncoherentParams()
    p.t0, p.dt, p.c = float(t0), float(dt), float(c)
    p.n_samples = int(n_samples)
    p.flags = _lib.SS_SPLIT_SIDES if split_sides else 0
    kid, pn, pa, pb, table = pattern_setup(pattern, dev)
    p.pattern, p.pat_n, p.pat_a = kid, pn, pa
    for i in range(4):
        p.pat_b[i] = pb[i]
    n_seg = (2 if split_sides else 1) * int(n_samples)
    T = traces.n
    if out is None:
        out = torch.empty((T, n_seg), dtype=torch.float64, device=dev)
    if dropped is None:
        dropped = torch.empty((T,), dtype=torch.float64, device=dev)
    pb_t = (torch.empty((T, facets.n), dtype=torch.int32, device=dev)
            if pair_bins else None)
    prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
    rc = _lib.load().ss_incoherent_power(
        prec, ctypes.byref(p), _ptr(facets.buf), facets.stride, facets.n,
        _ptr(traces.buf), traces.stride, T, _ptr(table), _ptr(out), _ptr(dropped),
        _ptr(pb_t), _stream_handle(stream, dev))
    _lib.check(rc, "ss_incoherent_power")
    if pair_bins:
        return out, dropped, pb_t, table
    return out, dropped, table


def is_horizontal(f):
    """Facets ``f`` form one horizontal plane: every normal (0, 0, +nz) and
    every centre at one z (exact comparisons)."""
    return bool(np.all(f.normals[:, :2] == 0.0) and np.all(f.normals[:, 2] > 0.0)
                and np.all(f.centers[:, 2] == f.centers[0, 2]))


def _iface_entry(f):
    """Lookup data of one interface: (table (rows, 6), descriptor row with
    TABOFF and SIGC unset, horizontal?)."""
    desc = np.zeros(IDESC["COUNT"], np.float64)
    ny, nx = f.grid_shape
    horizontal = is_horizontal(f)
    if horizontal:
        # every cell gives the same plane: one cell, fit (row, col) = 0
        ny = nx = 1
        tab = np.concatenate([f.centers[:1], f.normals[:1]], axis=1)
        coef = np.zeros((3, 2))
    else:
        tab = np.concatenate([f.centers.reshape(ny * nx, 3),
                              f.normals.reshape(ny * nx, 3)], axis=1)
        a = np.column_stack([f.centers[:, 0], f.centers[:, 1], np.ones(len(f.centers))])
        coef, *_ = np.linalg.lstsq(a, f.cell.astype(np.float64), rcond=None)
    w = (f.areas / f.areas.sum())[:, None]
    mp = (f.centers * w).sum(0)
    mn = (f.normals * w).sum(0)
    mn = mn / np.linalg.norm(mn)
    desc[IDESC["NY"]] = ny
    desc[IDESC["NX"]] = nx
    desc[IDESC["COEF"]:IDESC["COEF"] + 6] = coef.ravel()
    desc[IDESC["MP"]:IDESC["MP"] + 3] = mp
    desc[IDESC["MN"]:IDESC["MN"] + 3] = mn
    return tab, desc, horizontal


def _pad_planes_z(z_surf, n_real, pad_to, z_platform):
    """``_pad_planes`` from the surface's highest point."""
    k = 0 if pad_to is None else int(pad_to) - n_real
    if k < 0:
        raise ValueError(f"pad_to={pad_to} < {n_real} crossed interfaces")
    if not k:
        return []
    gap = float(z_platform) - z_surf
    if gap <= 0:
        raise ValueError("joint refraction padding needs the platform above the "
                         f"surface (gap {gap:.1f} m)")
    return list(z_surf + gap * (np.arange(k, 0, -1) / (k + 1.0)))


def _pad_planes(crossed, pad_to, z_platform):
    """z of the index-matched no-op planes the joint solver pads a stack
    with, spread through the air gap above the surface, top-down
    (``kernels/multilayer.py:286-330``)."""
    k = 0 if pad_to is None else int(pad_to) - len(crossed)
    if k < 0:
        raise ValueError(f"pad_to={pad_to} < {len(crossed)} crossed interfaces")
    if not k:
        return []
    z_surf = float(crossed[0].centers[:, 2].max())
    gap = float(z_platform) - z_surf
    if gap <= 0:
        raise ValueError("joint refraction padding needs the platform above the "
                         f"surface (gap {gap:.1f} m)")
    return list(z_surf + gap * (np.arange(k, 0, -1) / (k + 1.0)))


class DeviceInterfaces:
    """Crossed-interface lookup data for the refracted kernel: per interface
    an (ny*nx, 6) float64 [centre, normal] table in grid order, the affine
    (x, y, 1) -> (row, col) fit and the area-weighted mean plane
    (``kernels/multilayer.py:269-283``, float64).  A horizontal plane is
    stored as one cell: every lookup on it returns the same plane, so no
    crossing depends on the cell (the float64 two-point solve reads a cell
    only through n . (x - o), which for the exact +z normal is x_z - z, so
    the sequential chain's result is bit-identical).  ``stacks`` builds the
    stacks of several target layers over one shared table."""

    def __init__(self, crossed, crossed_sigma=None, device=None, pad_to=None,
                 z_platform=None):
        built = DeviceInterfaces.stacks([crossed], [crossed_sigma], [pad_to], z_platform,
                                        device)[0]
        self.__dict__.update(built.__dict__)

    @staticmethod
    def lattice_stacks(stacks, built, table, pad_row0, pad_tos=None, sigmas=None,
                       z_platform=None, device=None):
        """One DeviceInterfaces per crossed stack of Lattices tessellated by
        ``build_lattices`` (``built``, ``table``): descriptors point into the
        shared table; pads take rows from ``pad_row0`` on."""
        device = _dev(device)
        sigmas = sigmas or [None] * len(stacks)
        pad_tos = pad_tos or [None] * len(stacks)
        row = pad_row0
        pad_rows = []
        out = []
        for crossed, sig, pad_to in zip(stacks, sigmas, pad_tos):
            n_real = len(crossed)
            if n_real < 1 or n_real > _lib.MAX_CROSSED:
                raise ValueError(f"need 1..{_lib.MAX_CROSSED} crossed interfaces, got {n_real}")
            # the surface's highest facet centre (the reference pads from it)
            z_top = built[id(crossed[0])].z_max()
            pads = _pad_planes_z(z_top, n_real, pad_to, z_platform)
            k = len(pads)
            desc = np.zeros((n_real + k, IDESC["COUNT"]), np.float64)
            for i, z in enumerate(pads):
                desc[i, IDESC["NY"]] = desc[i, IDESC["NX"]] = 1
                desc[i, IDESC["MP"] + 2] = z
                desc[i, IDESC["MN"] + 2] = 1.0
                desc[i, IDESC["TABOFF"]] = row
                pad_rows.append((row, z))
                row += 1
            horizontal = []
            for i0, lat in enumerate(crossed):
                b = built[id(lat)]
                desc[i0 + k] = b.desc_row()
                desc[i0 + k, IDESC["SIGC"]] = 0.0 if sig is None else sig[i0]
                desc[i0 + k, IDESC["TABOFF"]] = b.table_off
                horizontal.append(b.horizontal)
            v = object.__new__(DeviceInterfaces)
            v.device, v.k_pad, v.n, v.has_sigma = device, k, desc.shape[0], sig is not None
            v.horizontal = horizontal
            v.flat_below = all(horizontal[1:])
            v.desc = torch.as_tensor(desc, device=device)
            v.tables = table
            out.append(v)
        if row > table.shape[0]:
            raise ValueError("shared table has too few pad rows")
        for r, z in pad_rows:
            table[r] = torch.tensor([0.0, 0.0, z, 0.0, 0.0, 1.0], dtype=torch.float64)
        return out

    @staticmethod
    def stacks(stacks, sigmas=None, pad_tos=None, z_platform=None, device=None):
        """One DeviceInterfaces per crossed stack, all over ONE device table:
        each distinct Facets object (by identity) is tabulated and uploaded
        once, the pads of every stack are rows of the same table."""
        device = _dev(device)
        sigmas = sigmas or [None] * len(stacks)
        pad_tos = pad_tos or [None] * len(stacks)
        entries, tabs, offs = {}, [], {}
        off = 0
        descs = []
        for crossed, sig, pad_to in zip(stacks, sigmas, pad_tos):
            n_real = len(crossed)
            if n_real < 1 or n_real > _lib.MAX_CROSSED:
                raise ValueError(f"need 1..{_lib.MAX_CROSSED} crossed interfaces, got {n_real}")
            pads = _pad_planes(crossed, pad_to, z_platform)
            k = len(pads)
            desc = np.zeros((n_real + k, IDESC["COUNT"]), np.float64)
            for i, z in enumerate(pads):
                desc[i, IDESC["NY"]] = desc[i, IDESC["NX"]] = 1
                desc[i, IDESC["MP"] + 2] = z
                desc[i, IDESC["MN"] + 2] = 1.0
                desc[i, IDESC["TABOFF"]] = off
                tabs.append(np.array([[0.0, 0.0, z, 0.0, 0.0, 1.0]]))
                off += 1
            horizontal = []
            for i0, f in enumerate(crossed):
                key = id(f)
                if key not in entries:
                    entries[key] = _iface_entry(f)
                    tabs.append(entries[key][0])
                    offs[key] = off
                    off += len(entries[key][0])
                _, row, hz = entries[key]
                desc[i0 + k] = row
                desc[i0 + k, IDESC["SIGC"]] = 0.0 if sig is None else sig[i0]
                desc[i0 + k, IDESC["TABOFF"]] = offs[key]
                horizontal.append(hz)
            descs.append((desc, k, horizontal, sig is not None))
        tables = torch.as_tensor(np.ascontiguousarray(np.concatenate(tabs)), device=device)
        out = []
        for desc, k, horizontal, has_sigma in descs:
            v = object.__new__(DeviceInterfaces)
            v.device, v.k_pad, v.n, v.has_sigma = device, k, desc.shape[0], has_sigma
            # per real interface: a horizontal plane (every cell normal
            # (0, 0, +nz) and one z), the condition of the flat-stack joint solve
            v.horizontal = horizontal
            v.flat_below = all(horizontal[1:])
            v.desc = torch.as_tensor(desc, device=device)
            v.tables = tables
            out.append(v)
        return out

    def nbytes(self):
        return (self.desc.numel() + self.tables.numel()) * 8

    def prefix(self, n):
        """View of the top ``n`` interfaces (a layer higher in the same
        stack): the descriptor rows' table offsets stay valid, so the tables
        are shared, not rebuilt.  Sequential (unpadded) stacks only."""
        if self.k_pad or not 1 <= n <= self.n:
            raise ValueError("prefix views need an unpadded stack and 1 <= n <= its depth")
        v = object.__new__(DeviceInterfaces)
        v.device, v.k_pad, v.n, v.has_sigma = self.device, 0, n, self.has_sigma
        v.horizontal = self.horizontal[:n]
        v.flat_below = all(v.horizontal[1:])
        v.desc = self.desc[:n]
        v.tables = self.tables
        return v


def refracted_params(target, ifaces, eps_leg, att_leg, *, mode, t0, dt, n_samples, c,
                     gamma=0.0, k0=None, split_sides=False, pattern=None, roughness=None,
                     taper_s=None, d_phi_area=False, n_exp=None, refraction="sequential",
                     joint_newton=6, joint_backtrack=4, extra_flags=0, flat_stack=False):
    """(RefractedParams, legs tensor, pattern table or None) for the
    refracted kernel over the (possibly padded) stack ``ifaces``;
    ``eps_leg``/``att_leg`` are the unpadded per-leg media.  ``extra_flags``
    is ORed into the flags (e.g. SS_REDO_ALL in tests).  ``flat_stack``
    (joint only): False (default) the reference's general joint solve,
    "auto" the flat-stack form (refraction code 2) when every real interface
    below the surface is a horizontal plane, True requires the flat one."""
    coherent = mode == "coherent"
    flags = _lib.SS_COHERENT if coherent else 0
    flags |= int(extra_flags)
    flags |= _lib.SS_SPLIT_SIDES if split_sides else 0
    flags |= _lib.SS_TAPER if taper_s is not None else 0
    flags |= _lib.SS_D_PHI_AREA if d_phi_area else 0
    flags |= _lib.SS_GAMMA_FACET if target.has_gamma else 0
    flags |= _lib.SS_CROSS_ROUGH if ifaces.has_sigma else 0
    flags |= _lib.SS_DIFFUSE if target.has_diffuse else 0
    p = _lib.RefractedParams()
    p.k0 = 0.0 if k0 is None else float(k0)
    p.gamma, p.t0, p.dt, p.c = float(gamma), float(t0), float(dt), float(c)
    p.taper_s = 0.0 if taper_s is None else float(taper_s)
    p.n_exp = 0.0 if n_exp is None else float(n_exp)
    p.n_samples = int(n_samples)
    p.n_crossed = ifaces.n
    if refraction not in ("sequential", "joint"):
        raise ValueError(f"unknown refraction solver {refraction!r}")
    if flat_stack not in ("auto", True, False):
        raise ValueError("flat_stack must be 'auto', True or False")
    if flat_stack is True and not ifaces.flat_below:
        raise ValueError("flat_stack=True needs horizontal planes below the surface")
    p.refraction = 0
    if refraction == "joint":
        p.refraction = 2 if (flat_stack is not False and ifaces.flat_below) else 1
    p.k_pad = ifaces.k_pad
    p.joint_newton, p.joint_backtrack = int(joint_newton), int(joint_backtrack)
    if roughness is not None:
        sigma, lc, n_terms, acf = roughness
        flags |= _lib.SS_ROUGH
        p.sigma, p.corr_len, p.n_terms = float(sigma), float(lc), int(n_terms)
        p.acf = _lib.ACFS[acf]
    kid, pn, pa, pb, table = pattern_setup(pattern, target.device)
    p.pattern, p.pat_n, p.pat_a = kid, pn, pa
    for i in range(4):
        p.pat_b[i] = pb[i]
    p.flags = flags
    eps = np.asarray(eps_leg, np.float64)
    att = np.asarray(att_leg, np.float64)
    if ifaces.k_pad:
        eps = np.concatenate([np.full(ifaces.k_pad, eps[0]), eps])
        att = np.concatenate([np.full(ifaces.k_pad, att[0]), att])
    legs = np.stack([np.sqrt(eps), eps, att])
    if legs.shape[1] != ifaces.n + 1:
        raise ValueError(f"need {ifaces.n + 1 - ifaces.k_pad} legs, got {len(eps_leg)}")
    legs_t = torch.as_tensor(np.ascontiguousarray(legs), device=target.device)
    return p, legs_t, table


def refracted_field(target, traces, ifaces, eps_leg, att_leg, *, mode, t0, dt,
                    n_samples, c, gamma=0.0, k0=None, precision="mixed",
                    split_sides=False, pattern=None, roughness=None,
                    taper_s=None, d_phi_area=False, n_exp=None, refraction="sequential",
                    joint_newton=6, joint_backtrack=4, out=None,
                    dropped=None, pair_bins=False, workspace=None, extra_flags=0,
                    stream=None, flat_stack=False, noalloc=False):
    """Enqueue the refracted kernel (sequential chain, joint solve or its
    flat-stack form, see ``refracted_params``) and,
    for the mixed tier, its float64 redo launch.  ``workspace``: a uint8
    device tensor of at least ``refracted_workspace_bytes`` (allocated here
    when None).  ``noalloc``: call ss_refracted_field_noalloc instead (the
    library's own per-device scratch array holds the redo record; no
    workspace).  Returns (out (T, n_seg, 2) float64, dropped (T,)
    [, pair_bins], keep-alive) device tensors."""
    dev = target.device
    p, legs_t, table = refracted_params(
        target, ifaces, eps_leg, att_leg, mode=mode, t0=t0, dt=dt, n_samples=n_samples,
        c=c, gamma=gamma, k0=k0, split_sides=split_sides, pattern=pattern,
        roughness=roughness, taper_s=taper_s, d_phi_area=d_phi_area, n_exp=n_exp,
        refraction=refraction, joint_newton=joint_newton, joint_backtrack=joint_backtrack,
        extra_flags=extra_flags, flat_stack=flat_stack)
    n_seg = (2 if split_sides else 1) * int(n_samples)
    T = traces.n
    if out is None:
        out = torch.empty((T, n_seg, 2), dtype=torch.float64, device=dev)
    if dropped is None:
        dropped = torch.empty((T,), dtype=torch.float64, device=dev)
    pb_t = (torch.empty((T, target.n), dtype=torch.int32, device=dev)
            if pair_bins else None)
    cst = constants(dev)
    prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
    if noalloc:
        rc = _lib.load().ss_refracted_field_noalloc(
            prec, ctypes.byref(p), _ptr(target.buf), target.stride, target.n,
            _ptr(traces.buf), traces.stride, T, _ptr(ifaces.desc), _ptr(ifaces.tables),
            _ptr(legs_t), _ptr(table), _ptr(cst), _ptr(out), _ptr(dropped), _ptr(pb_t),
            _stream_handle(stream, dev))
        _lib.check(rc, "ss_refracted_field_noalloc")
    else:
        need = refracted_workspace_bytes(precision, target.n, T)
        if workspace is None and need > 0:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
        ws_bytes = 0 if workspace is None else workspace.numel()
        rc = _lib.load().ss_refracted_field(
            prec, ctypes.byref(p), _ptr(target.buf), target.stride, target.n,
            _ptr(traces.buf), traces.stride, T, _ptr(ifaces.desc), _ptr(ifaces.tables),
            _ptr(legs_t), _ptr(table), _ptr(cst), _ptr(out), _ptr(dropped), _ptr(pb_t),
            _ptr(workspace), ws_bytes, _stream_handle(stream, dev))
        _lib.check(rc, "ss_refracted_field")
    keep = (legs_t, table, workspace)
    if pair_bins:
        return out, dropped, pb_t, keep
    return out, dropped, keep


def trace_diagnostics(fac, traces, *, t0, dt, n_samples, c, stream=None):
    """Per-trace nadir facet and first return over ``fac``
    (``ss_trace_diagnostics``): dict of device tensors nadir_index,
    nadir_z, first_index, first_twtt, first_xyz (indices in the caller's
    facet order)."""
    dev = fac.device
    T = traces.n
    idx = getattr(fac, "_index_t", None)
    if idx is None:
        idx = fac._index_t = torch.as_tensor(np.asarray(fac.order, np.int64), device=dev)
    out = {"nadir_index": torch.empty(T, dtype=torch.int64, device=dev),
           "nadir_z": torch.empty(T, dtype=torch.float64, device=dev),
           "first_index": torch.empty(T, dtype=torch.int64, device=dev),
           "first_twtt": torch.empty(T, dtype=torch.float64, device=dev),
           "first_xyz": torch.empty((T, 3), dtype=torch.float64, device=dev)}
    rc = _lib.load().ss_trace_diagnostics(
        _ptr(fac.buf), fac.stride, fac.n, _ptr(idx), _ptr(traces.buf), traces.stride, T,
        float(t0), float(dt), int(n_samples), float(c), _ptr(out["nadir_index"]),
        _ptr(out["nadir_z"]), _ptr(out["first_index"]), _ptr(out["first_twtt"]),
        _ptr(out["first_xyz"]), _stream_handle(stream, dev))
    _lib.check(rc, "ss_trace_diagnostics")
    return out


def refracted_workspace_bytes(precision, n_facets, n_traces):
    """Device workspace bytes of one refracted launch (0 for fp64)."""
    prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
    return int(_lib.load().ss_refracted_workspace_bytes(prec, int(n_facets), int(n_traces)))


class PulseSpectrum:
    """A pulse prepared for range compression on ``device``: its taps as a
    device tensor and, when the FFT path applies to rows of ``n_samples``,
    its spectrum (computed once on the device)."""

    def __init__(self, pulse, m, n_samples, device, method="auto", stream=None):
        p = np.asarray(pulse)
        self.m = int(m)
        self.n_samples = int(n_samples)
        self.taps = torch.as_tensor(
            np.stack([p.real, p.imag], -1).astype(np.float64) if np.iscomplexobj(p) else
            np.stack([p, np.zeros_like(p)], -1).astype(np.float64),
            device=_dev(device)).contiguous()
        lib = _lib.load()
        self.nfft = 0 if method == "direct" else int(lib.ss_compress_fft_size(self.n_samples,
                                                                              self.m))
        if method == "fft" and self.nfft == 0:
            raise ValueError("pulse too long for the FFT path")
        self.spectrum = None
        if self.nfft:
            self.spectrum = torch.empty((self.nfft, 2), dtype=torch.float64,
                                        device=self.taps.device)
            rc = lib.ss_pulse_spectrum(_ptr(self.taps), self.m, self.nfft,
                                       _ptr(self.spectrum),
                                       _stream_handle(stream, self.taps.device))
            _lib.check(rc, "ss_pulse_spectrum")


def range_compress(field, pulse, m=None, out=None, stream=None, method="auto"):
    """'Same' convolution of every row of a (..., n, 2) float64 device tensor
    (complex interleaved) with a complex pulse of 2m+1 taps (or a prepared
    PulseSpectrum): FFT overlap-save when the pulse fits the LDS transform
    (``method`` "auto"/"fft"), else direct summation ("direct").  Returns
    (out, the PulseSpectrum used)."""
    n = field.shape[-2]
    rows = field.numel() // (2 * n)
    ps = pulse if isinstance(pulse, PulseSpectrum) else PulseSpectrum(
        pulse, m, n, field.device, method=method, stream=stream)
    if ps.n_samples != n:
        raise ValueError(f"spectrum prepared for {ps.n_samples} samples, rows have {n}")
    src = field.contiguous()
    if out is None:
        out = torch.empty_like(src)
    lib = _lib.load()
    use_fft = ps.nfft > 0 and method != "direct"
    for r0 in range(0, rows, 65535):
        nr = min(65535, rows - r0)
        off = r0 * n * 2
        if use_fft:
            rc = lib.ss_range_compress_fft(ctypes.c_void_p(src.data_ptr() + off * 8), nr, n,
                                           _ptr(ps.spectrum), ps.m, ps.nfft,
                                           ctypes.c_void_p(out.data_ptr() + off * 8),
                                           _stream_handle(stream, field.device))
            _lib.check(rc, "ss_range_compress_fft")
        else:
            rc = lib.ss_range_compress(ctypes.c_void_p(src.data_ptr() + off * 8), nr, n,
                                       _ptr(ps.taps), ps.m,
                                       ctypes.c_void_p(out.data_ptr() + off * 8),
                                       _stream_handle(stream, field.device))
            _lib.check(rc, "ss_range_compress")
    return out, ps


def slow_time_boxcar(x, window, stride, n_out, scale=1.0, stream=None):
    """``ss_slow_time_boxcar`` on a (rows, ...) float64 device tensor
    (trailing dims flattened into columns); returns (n_out, ...) float64."""
    x = x.contiguous()
    rows = x.shape[0]
    cols = x[0].numel() if rows else 0
    out = torch.empty((n_out,) + tuple(x.shape[1:]), dtype=torch.float64, device=x.device)
    rc = _lib.load().ss_slow_time_boxcar(_ptr(x), rows, cols, int(window), int(stride),
                                         int(n_out), float(scale), _ptr(out),
                                         _stream_handle(stream, x.device))
    _lib.check(rc, "ss_slow_time_boxcar")
    return out


def focused_sar(field, twtt, s, *, c, k, aperture_m, window="hann", precision="fp64",
                stream=None):
    """``ss_focused_sar`` on a (T, nb, 2) float64 device field; twtt (nb,)
    and s (T,) device tensors.  Returns (T, nb, 2) float64."""
    if window not in ("hann", "none", "rect", None):
        raise ValueError(f"unknown window {window!r} (use 'hann' or 'none')")
    field = field.contiguous()
    T, nb = field.shape[0], field.shape[1]
    out = torch.empty_like(field)
    prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
    rc = _lib.load().ss_focused_sar(prec, _ptr(field), T, nb, _ptr(twtt.contiguous()),
                                    _ptr(s.contiguous()), float(c), float(k),
                                    float(aperture_m), 1 if window == "hann" else 0,
                                    _ptr(out), _stream_handle(stream, field.device))
    _lib.check(rc, "ss_focused_sar")
    return out

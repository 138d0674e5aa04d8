"""Batched persistent launches over uneven requests (``ss_batch_field``).

A ``BatchRunner`` owns the device buffers the C entry point needs (the job
table and the workspace: work counter + float64 redo queue), grown only when
a larger batch arrives, so a steady launch allocates nothing.  Every
layer of every scene becomes one ``ss_job``; one persistent launch drains all
(job, trace tile, facet tile) work items; range compression follows per scene.

``uneven_workloads`` draws the 64-scene mix: facets log-uniform in
[1e4, 2e6] per interface, traces uniform in [50, 2000], 1-4 layers, centre
frequency 30-330 MHz with 10-30 MHz bandwidth, per-scene seeds; subsurface
layers follow the reference's ``simulate()``: the sequential chain for the
first one under the surface, the joint solve (``SimConfig.refraction``'s
default) below it, padded to the reference's power-of-two buckets
(``simulate.py:253-260, 357``; with at most 4 interfaces no pads arise).
``scale`` shrinks facet and trace counts for quick runs (1.0 = the stated
distribution).
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import device as dv
from .physics import C, fresnel_normal
from .scene import LatticeFacets, SeededLattice, grid_facets, straight_track
from .waveform import Waveform
from .workloads import Layer, Workload, _pad_bucket


class BatchRunner:
    def __init__(self, device=None, max_jobs=1024):
        self.device = dv._dev(device)
        self.size = int(_lib.load().ss_sizeof_job())
        self.jobs_dev = torch.empty(max_jobs * self.size, dtype=torch.uint8, device=self.device)
        self.workspace = torch.empty(256, dtype=torch.uint8, device=self.device)
        self.max_jobs = max_jobs
        self._host = None

    def launch(self, jobs, precision="mixed", trace_tile=8, n_blocks=0, stream=None):
        """Enqueue one persistent launch over ``jobs`` (list of _lib.Job)."""
        n = len(jobs)
        if n > self.max_jobs:
            raise ValueError(f"{n} jobs > {self.max_jobs}")
        arr = (_lib.Job * max(n, 1))(*jobs)
        self._host = arr            # the async copy reads it; keep it alive
        prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
        lib = _lib.load()
        need = int(lib.ss_batch_workspace_bytes(prec, arr, n, int(trace_tile)))
        if need < 0:
            raise ValueError(f"trace_tile must be 4, 8 or 16, got {trace_tile}")
        if need > self.workspace.numel():
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        rc = lib.ss_batch_field(prec, arr, n, ctypes.c_void_p(self.jobs_dev.data_ptr()),
                                ctypes.c_void_p(self.workspace.data_ptr()),
                                self.workspace.numel(), int(trace_tile), int(n_blocks),
                                dv._stream_handle(stream, self.device))
        _lib.check(rc, "ss_batch_field")

    def run_scenes(self, scenes, precision="mixed", trace_tile=8, n_blocks=0, stream=None,
                   compress=True):
        jobs, keep = [], []
        for rs in scenes:
            j, k = rs.jobs(precision)
            jobs += j
            keep += k
        self.launch(jobs, precision, trace_tile, n_blocks, stream)
        if compress:
            for rs in scenes:
                rs.compress(stream)
        return keep


def pack_info(jobs, trace_tile):
    """Validate ``jobs`` and return (surface items, refracted items, feature
    class) at ``trace_tile`` (host only; ``ss_batch_pack`` into scratch)."""
    n = len(jobs)
    lib = _lib.load()
    arr = (_lib.Job * max(n, 1))(*jobs)
    items = (ctypes.c_int64 * 2)()
    cls = ctypes.c_int()
    scratch = np.zeros(int(lib.ss_batch_table_bytes(max(min(n, 1024), 1))), np.uint8)
    rc = lib.ss_batch_pack(arr, n, int(trace_tile), scratch.ctypes.data, scratch.size, items,
                           ctypes.byref(cls))
    _lib.check(rc, "ss_batch_pack")
    return int(items[0]), int(items[1]), cls.value


class StepGraph:
    """One batched step captured in a HIP graph (``ss_batch_step``): a copy
    of a pinned host job table into a device table, then the step's kernels,
    all reading the table in device memory.  ``launch(jobs)`` repacks the
    pinned table (``ss_batch_pack``) and replays the graph, so a step costs
    one graph launch whatever the number of jobs.  A graph is specific to
    (precision, trace tile, feature class) and holds at most ``max_jobs``
    jobs and ``max_items`` refracted work items.  The pinned table is reused
    by the next launch: wait for ``event`` (recorded after each launch)
    before launching the same StepGraph again.  ``capture=False`` (or a host
    device) enqueues the same launches without a graph."""

    def __init__(self, device, precision, trace_tile, feature_class, max_jobs=1024,
                 max_items=1 << 16, n_blocks=0, capture=True):
        self.device = dv._dev(device)
        cuda = self.device.type == "cuda"
        self.precision = precision
        self.prec = _lib.SS_PREC_FP64 if precision == "fp64" else _lib.SS_PREC_MIXED
        self.tt, self.cls = int(trace_tile), int(feature_class)
        self.max_jobs, self.max_items, self.n_blocks = int(max_jobs), int(max_items), int(n_blocks)
        lib = _lib.load()
        tb = int(lib.ss_batch_table_bytes(self.max_jobs))
        if tb < 0:
            raise ValueError(f"max_jobs out of range: {max_jobs}")
        self.table_host = torch.zeros(tb, dtype=torch.uint8, pin_memory=cuda)
        self.table_dev = torch.zeros(tb, dtype=torch.uint8, device=self.device)
        ws = int(lib.ss_batch_step_workspace_bytes(self.prec, self.max_items, self.tt))
        self.workspace = torch.empty(max(ws, 256), dtype=torch.uint8, device=self.device)
        self._arr = None
        self.event = None
        self.graph = None
        if capture and cuda:
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._enqueue()

    def _enqueue(self):
        self.table_dev.copy_(self.table_host, non_blocking=True)
        rc = _lib.load().ss_batch_step(
            self.prec, ctypes.c_void_p(self.table_dev.data_ptr()),
            ctypes.c_void_p(self.workspace.data_ptr()), self.workspace.numel(),
            self.max_items, self.tt, self.n_blocks, self.cls,
            dv._stream_handle(None, self.device))
        _lib.check(rc, "ss_batch_step")

    def fits(self, jobs):
        """(fits, (surface, refracted) items, class) of ``jobs`` here."""
        si, ri, cls = pack_info(jobs, self.tt)
        ok = len(jobs) <= self.max_jobs and ri <= self.max_items and cls <= self.cls
        return ok, (si, ri), cls

    def launch(self, jobs):
        n = len(jobs)
        if n > self.max_jobs:
            raise ValueError(f"{n} jobs > {self.max_jobs}")
        if self.event is not None:
            self.event.synchronize()          # the previous copy has read the table
        arr = (_lib.Job * max(n, 1))(*jobs)
        items = (ctypes.c_int64 * 2)()
        cls = ctypes.c_int()
        rc = _lib.load().ss_batch_pack(arr, n, self.tt,
                                       ctypes.c_void_p(self.table_host.data_ptr()),
                                       self.table_host.numel(), items, ctypes.byref(cls))
        _lib.check(rc, "ss_batch_pack")
        if items[1] > self.max_items or cls.value > self.cls:
            raise ValueError(f"step exceeds the graph: {items[1]} refracted items "
                             f"(max {self.max_items}), class {cls.value} (graph {self.cls})")
        self._arr = arr
        if self.graph is not None:
            self.graph.replay()
        else:
            self._enqueue()
        if self.device.type == "cuda":
            self.event = torch.cuda.Event()
            self.event.record()


def draw_scene(rng, key, scale=1.0, max_facets=None):
    """One scene of the uneven mix: (layers, deepest mean depth).  Facets
    log-uniform in [1e4, 2e6] x scale per interface, 1-4 interfaces: a
    surface and internal layers 150-800 m apart (eps 3.15, 10 dB/km ice;
    internal layers -40 dB, the last one a bed over eps 8).  Interface j is
    the seeded lattice of seed (*key, j)."""
    F = math.exp(rng.uniform(math.log(1e4), math.log(2e6))) * scale
    if max_facets:
        F = min(F, max_facets)
    L = int(rng.integers(1, 5))
    side = max(8, int(round(math.sqrt(F))))
    dx = 25.0
    lat = SeededLattice(side, side, dx, sigma=2.0, corr_len=200.0, seed=(*key, 0))
    surf = LatticeFacets(lat)
    eps_ice = 3.15
    layers = [Layer("surface", surf, fresnel_normal(1.0, eps_ice), lattice=lat)]
    crossed = [surf]
    eps = [1.0, eps_ice]
    att = [0.0, 10.0]
    depth = 0.0
    for j in range(1, L):
        depth += rng.uniform(150.0, 800.0)
        lat = SeededLattice(side, side, dx, sigma=20.0, corr_len=1000.0, seed=(*key, j),
                            mean=-depth)
        f = LatticeFacets(lat)
        eps_below = eps_ice if j < L - 1 else 8.0
        joint = j > 1
        layers.append(Layer(f"layer{j}", f, fresnel_normal(eps[-1], eps_below)
                            if j == L - 1 else -10 ** (-40 / 20), crossed=list(crossed),
                            eps_leg=list(eps), att_leg=list(att), lattice=lat,
                            refraction="joint" if joint else "sequential",
                            pad_to=_pad_bucket(j, L - 1) if joint else None))
        crossed.append(f)
        eps = eps + [eps_ice]
        att = att + [10.0]
    return layers, depth


def draw_instrument(rng, name, layers, depth, scale=1.0, n_samples=None, cross_offset=0.0):
    """A Workload observing ``layers`` with a freshly drawn instrument and
    trajectory: traces uniform in [50, 2000] x scale at 25 m along a track
    shifted ``cross_offset`` m across, centre frequency 30-330 MHz,
    bandwidth 10-30 MHz, sampling 2.5 B, a window reaching past the deepest
    layer."""
    T = int(max(4, round(rng.uniform(50, 2000) * scale)))
    f0 = rng.uniform(30e6, 330e6)
    B = rng.uniform(10e6, 30e6)
    alt = 20000.0
    dt = 1.0 / (2.5 * B)
    n = n_samples or int(2 ** math.ceil(math.log2(2.0 * (depth * 1.78 + 2000.0) / C / dt)))
    track = straight_track(T, 25.0, alt, cross_offset=cross_offset)
    F = len(layers[0].facets)
    return Workload(
        name=name, track=track, layers=layers, k0=2 * math.pi * f0 / C,
        t0=2.0 * (alt - 100.0) / C, dt=dt, n_samples=n, c=C,
        waveform=Waveform(kind="chirp", bandwidth=B, pulse_length=10e-6),
        meta=dict(f0=f0, B=B, F=F, T=T, L=len(layers)))


def uneven_workloads(n_scenes=64, seed=4, scale=1.0, n_samples=None, max_facets=None):
    """The uneven 64-scene mix (module docstring) as a list of Workloads."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        layers, depth = draw_scene(rng, (seed, s), scale, max_facets)
        out.append(draw_instrument(rng, f"scene{s}", layers, depth, scale, n_samples))
    return out

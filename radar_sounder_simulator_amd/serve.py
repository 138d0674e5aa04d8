"""Request loop with continuous batching over resident scenes.

A ``Server`` keeps scenes -- every layer's facets and crossed-interface
tables -- resident in HBM and admits requests by their HBM footprint
against a byte budget (by default the device's memory less a reserve).
Admitted requests run together: every step is one persistent batched
launch (``ss_batch_field``) over trace chunks of all runnable requests,
sized to a pair budget shared evenly between them, so a request that
arrives while others are on the GPU joins the next step rather than
waiting for them to finish, and its work overlaps theirs on the device.
A request that brings a new scene is admitted with the scene's footprint
reserved, and the scene is prepared and uploaded by a worker thread on its
own stream while the device keeps running the other requests; the request
joins the steps once the upload's event has completed.  A request's last
step is followed by its range compression and an asynchronous copy to
pinned host memory; its latency runs from arrival to the end of that copy.  Memory is returned when the copy completes; scenes
no request uses are evicted least-recently-used first when an admission
needs their room.

Each trace of a request is computed by exactly one work item set per layer
whichever step carries it, so a result is the same alone and in a mix up to
the order of the floating-point sums.

The reference has no request loop: its requests are one ``simulate()``
call after another (``simulate.py:172``).  This module is the serving
stream's host side; the kernels are the batched ones of ``batch.py``.
"""

import collections
import contextlib
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import device as dv
from .batch import BatchRunner, StepGraph, draw_instrument, draw_scene, pack_info
from .kernels import _reference_range
from .waveform import compressed_pulse


@dataclass
class Request:
    """One simulation request: the scene it observes (``scene_id``; the
    workload's layers are that scene's), the instrument and trajectory
    (the rest of ``workload``) and its arrival time (s from stream start)."""
    req_id: int
    scene_id: str
    workload: object
    arrival: float = 0.0


@dataclass
class _Pending:
    req: Request
    admitted: float
    bytes: int
    uploaded: bool


@dataclass
class Result:
    req_id: int
    scene_id: str
    field: np.ndarray          # (L, T, n_seg) complex128, range-compressed
    dropped: np.ndarray        # (L, T)
    pairs: int
    arrival: float
    admitted: float
    done: float
    uploaded: bool
    steps: int = 1             # batched steps that carried this request

    @property
    def latency(self):
        return self.done - self.arrival


class DeviceScene:
    """A scene's layers resident in HBM (facets + crossed-interface tables).
    When every layer's crossed interfaces are the top of one stack (the
    layered scenes here), the stack's tables are built once and each layer
    uses a prefix view of them."""

    def __init__(self, layers, device):
        self.layers = []
        self.center_mean = None
        for L in layers:
            if L.crossed and L.pad_to and L.pad_to > len(L.crossed):
                # pads sit between the surface and each request's platform:
                # a resident stack shared by requests carries none
                raise ValueError(f"layer {L.name}: padded joint stacks are not served "
                                 "(pad_to > crossed interfaces)")
        lat_of = {id(L.facets): L.lattice for L in layers if L.lattice is not None}
        if (all(L.lattice is not None for L in layers)
                and all(id(c) in lat_of for L in layers for c in L.crossed)):
            self._build_on_device(layers, lat_of, device)
            return
        deepest = max((L.crossed for L in layers if L.crossed), key=len, default=None)
        stack = None
        if deepest and all(not L.crossed or all(a is b for a, b in zip(L.crossed, deepest))
                           for L in layers):
            stack = dv.DeviceInterfaces(deepest, device=device)
        self.stack = stack
        for L in layers:
            ph = None if L.roughness is None else L.roughness[2]
            g = np.asarray(L.gamma, np.float64)
            fac = dv.DeviceFacets.from_facets(L.facets, gamma=g if g.ndim else None,
                                              phasors=ph, device=device)
            ifc = None
            if L.crossed:
                ifc = (stack.prefix(len(L.crossed)) if stack is not None else
                       dv.DeviceInterfaces(L.crossed, device=device))
            self.layers.append((L, fac, ifc))
        self.refs = 0

    def _build_on_device(self, layers, lat_of, device):
        """Every layer a Lattice: tessellated in HBM (ss_facets_from_heights),
        the heights the only per-facet upload; one shared lookup table."""
        lats = [L.lattice for L in layers]
        crossed = {id(lat_of[id(c)]) for L in layers for c in L.crossed}
        gammas, phasors = {}, {}
        for L in layers:
            g = np.asarray(L.gamma, np.float64)
            if g.ndim:
                gammas[id(L.lattice)] = g
            if L.roughness is not None:
                phasors[id(L.lattice)] = L.roughness[2]
        built, table = dv.build_lattices(lats, crossed, device=device, gammas=gammas,
                                         phasors=phasors)
        refr = [L for L in layers if L.crossed]
        ifcs = iter(dv.DeviceInterfaces.lattice_stacks(
            [[lat_of[id(c)] for c in L.crossed] for L in refr], built, table,
            pad_row0=table.shape[0], device=device) if refr else [])
        self.stack = None
        self.table = table
        for L in layers:
            self.layers.append((L, built[id(L.lattice)].fac, next(ifcs) if L.crossed else None))
        self.center_mean = built[id(lats[0])].center_mean
        self.refs = 0

    def nbytes(self):
        b = sum(f.nbytes() for _, f, _ in self.layers)
        if getattr(self, "table", None) is not None:
            return b + self.table.numel() * 8 + sum(i.desc.numel() * 8
                                                    for _, _, i in self.layers if i is not None)
        if self.stack is not None:
            return b + self.stack.nbytes()
        return b + sum(i.nbytes() for _, _, i in self.layers if i is not None)

    @staticmethod
    def estimate(layers):
        """HBM bytes DeviceScene(layers) will take (facet SoA rows + tables;
        an upper bound when the layers share one stack)."""
        b = 0
        for L in layers:
            b += _lib.FROW["COUNT"] * len(L.facets) * 8
            for c in (L.crossed or []):
                b += (len(c) * 6 + 16) * 8
        return b


def _range_to(positions, m):
    """Per-trace |p - m| in float64 (the surface reference range from a
    device-side centre mean)."""
    d = np.asarray(positions, np.float64) - np.asarray(m, np.float64)
    return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])


class ActiveRequest:
    """An admitted request: its traces, output buffers, per-layer job
    templates and how far each layer's traces have been launched."""

    def __init__(self, req, scene, device, precision, admitted):
        self.req = req
        self.scene = scene
        self.admitted = admitted
        wl = req.workload
        tr = wl.track
        pv = None if wl.pattern is None else wl.pattern[1]
        surf = wl.layers[0].facets
        self.traces = dv.DeviceTraces(tr.positions, tr.u_ct,
                                      _reference_range(tr.positions, surf.centers)
                                      if scene.center_mean is None else
                                      _range_to(tr.positions, scene.center_mean), pv,
                                      device=device)
        self.T = T = tr.positions.shape[0]
        self.n_seg = n_seg = (2 if wl.split_sides else 1) * wl.n_samples
        nl = len(scene.layers)
        self.raw = torch.empty((nl, T, n_seg, 2), dtype=torch.float64, device=device)
        self.out = torch.empty_like(self.raw)
        self.dropped = torch.empty((nl, T), dtype=torch.float64, device=device)
        self.pulse = None
        if wl.waveform is not None and wl.waveform.kind != "delta":
            p, m = compressed_pulse(wl.waveform.bandwidth, wl.waveform.pulse_length,
                                    wl.dt, wl.waveform.window,
                                    wl.waveform.construction, wl.waveform.tukey_alpha)
            self.pulse = dv.PulseSpectrum(p, m, wl.n_samples, device)
        self.keep = []
        self.templates = []
        cst = dv.constants(device)
        for L, fac, ifc in scene.layers:
            jb = _lib.Job()
            jb.facet_stride, jb.n_facets = fac.stride, fac.n
            jb.trace_stride = self.traces.stride
            jb.facets = fac.buf.data_ptr()
            jb.constants = cst.data_ptr()
            g = 0.0 if fac.has_gamma else L.gamma
            rough = (None if L.roughness is None else
                     (L.roughness[0], L.roughness[1], L.roughness[3],
                      L.roughness[4] if len(L.roughness) > 4 else "gaussian"))
            if ifc is None:
                p, table = dv.surface_params(
                    fac, k=wl.k0, gamma=g, t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples,
                    c=wl.c, split_sides=wl.split_sides, pattern=wl.pattern, roughness=rough,
                    taper_s=wl.taper_s, d_phi_area=wl.d_phi_area)
                jb.kind = 0
                jb.surface = p
            else:
                p, legs, table = dv.refracted_params(
                    fac, ifc, L.eps_leg, L.att_leg, mode="coherent", t0=wl.t0, dt=wl.dt,
                    n_samples=wl.n_samples, c=wl.c, gamma=g, k0=wl.k0,
                    split_sides=wl.split_sides, pattern=wl.pattern, roughness=rough,
                    taper_s=wl.taper_s, d_phi_area=wl.d_phi_area, refraction=L.refraction,
                    flat_stack=L.flat_stack)
                jb.kind = 1
                jb.refracted = p
                jb.iface_desc, jb.tables = ifc.desc.data_ptr(), ifc.tables.data_ptr()
                jb.legs = legs.data_ptr()
                self.keep.append(legs)
            jb.pattern_table = 0 if table is None else table.data_ptr()
            self.keep.append(table)
            self.templates.append(jb)
        self.cursor = [0] * nl
        self.n_steps = 0
        self.layer_facets = [fac.n for _, fac, _ in scene.layers]
        self.pairs = T * sum(self.layer_facets)
        self.event = None
        self.host_out = None
        self.host_drop = None

    @staticmethod
    def estimate(req):
        """HBM bytes an admitted request holds (traces, raw + compressed
        fields, dropped power)."""
        wl = req.workload
        T = wl.track.positions.shape[0]
        n_seg = (2 if wl.split_sides else 1) * wl.n_samples
        nl = len(wl.layers)
        return _lib.TROW["COUNT"] * T * 8 + 2 * nl * T * n_seg * 16 + nl * T * 8

    def remaining(self):
        return sum(F * (self.T - c) for F, c in zip(self.layer_facets, self.cursor))

    def take(self, budget, tt):
        """Jobs for up to ``budget`` pairs of this request's next traces
        (whole trace tiles, at least one tile), layer by layer."""
        jobs = []
        T = self.T
        self.n_steps += 1
        for j, F in enumerate(self.layer_facets):
            if budget <= 0:
                break
            c = self.cursor[j]
            if c >= T:
                continue
            n = max(tt, int(budget // max(F, 1)) // tt * tt)
            n = min(n, T - c)
            jobs.append(self._chunk(j, c, c + n))
            self.cursor[j] = c + n
            budget -= n * F
        return jobs

    def _chunk(self, j, t0, t1):
        jb = _lib.Job.from_buffer_copy(self.templates[j])
        jb.traces = self.traces.buf.data_ptr() + 8 * t0
        jb.n_traces = t1 - t0
        jb.field = self.raw.data_ptr() + 16 * (j * self.T + t0) * self.n_seg
        jb.dropped = self.dropped.data_ptr() + 8 * (j * self.T + t0)
        return jb


class CaptureGate:
    """Keeps graph capture and the upload workers' device work apart.

    A stream capture in global mode (torch's default) is invalidated by an
    allocation, a synchronous copy or a synchronisation that ANY thread of
    the process makes while it runs, so a scene upload on a worker thread
    overlapping the capture of a step breaks the capture.  Workers enter
    ``work()`` around their device calls (any number at once); ``capture()``
    waits until none is inside and holds new ones off until it ends."""

    def __init__(self):
        self._cv = threading.Condition()
        self._workers = 0
        self._capturing = False

    @contextlib.contextmanager
    def work(self):
        with self._cv:
            self._cv.wait_for(lambda: not self._capturing)
            self._workers += 1
        try:
            yield
        finally:
            with self._cv:
                self._workers -= 1
                self._cv.notify_all()

    @contextlib.contextmanager
    def capture(self):
        with self._cv:
            self._cv.wait_for(lambda: self._workers == 0 and not self._capturing)
            self._capturing = True
        try:
            yield
        finally:
            with self._cv:
                self._capturing = False
                self._cv.notify_all()


class Server:
    """Continuous-batching request loop on one device (see module docstring).

    ``budget_bytes``: HBM the server may hold for scenes and requests
    (default: total device memory less 8 GiB of reserve).  ``step_pairs``:
    pair budget of one batched step.  ``trace_tile``/``n_blocks``: the
    batched launch's tile and grid (autotuned values from ``autotune``)."""

    def __init__(self, device=None, precision="mixed", budget_bytes=None, step_pairs=4e8,
                 trace_tile=8, n_blocks=0, max_inflight=2, max_jobs=1024, upload_workers=2,
                 graphs=True):
        self.device = dv._dev(device)
        self.cuda = self.device.type == "cuda"
        self.precision = precision
        if budget_bytes is None:
            if not self.cuda:
                raise ValueError("budget_bytes is required off the GPU")
            total = torch.cuda.get_device_properties(self.device).total_memory
            budget_bytes = total - (8 << 30)
        self.budget = int(budget_bytes)
        self.step_pairs = float(step_pairs)
        self.tt = int(trace_tile)
        self.n_blocks = int(n_blocks)
        self.max_inflight = int(max_inflight)
        self.max_jobs = int(max_jobs)
        self.runner = BatchRunner(device=self.device, max_jobs=max_jobs)
        self.scenes = collections.OrderedDict()    # scene_id -> DeviceScene (LRU order)
        self.scene_bytes = {}
        self.uploads = {}                           # scene_id -> (future, reserved bytes)
        self.pool = ThreadPoolExecutor(max_workers=upload_workers)
        self.gate = CaptureGate()
        self.upload_stream = torch.cuda.Stream(self.device) if self.cuda else None
        # captured steps: two StepGraphs per feature class, used alternately
        # so a step's pinned table is never rewritten while in flight
        self.graphs = graphs and self.cuda
        self._graph_sets = {}
        self._graph_turn = 0
        self.graph_steps = 0
        # captured-step capacity in (trace tile, 256-facet) work items: four
        # times a step's share, at most 1 << 22 (a mixed-tier workspace of
        # about 1 GiB per graph); a larger step runs as an eager launch
        self.max_items = int(min(max(1 << 16, 4 * self.step_pairs / (256 * self.tt)), 1 << 22))
        self.used = 0
        self.peak = 0
        self.steps = 0

    # ------------------------------------------------------------ residency
    def add_scene(self, scene_id, layers):
        """Upload a scene and keep it resident (counted against the budget)."""
        if scene_id in self.scenes:
            self.scenes.move_to_end(scene_id)
            return self.scenes[scene_id]
        need = DeviceScene.estimate(layers)
        if not self._make_room(need):
            raise MemoryError(f"scene {scene_id} ({need} B) does not fit the budget")
        sc = DeviceScene(layers, self.device)
        self.scenes[scene_id] = sc
        self.scene_bytes[scene_id] = sc.nbytes()
        self._account(sc.nbytes())
        return sc

    def _account(self, delta):
        self.used += delta
        self.peak = max(self.peak, self.used)

    def _make_room(self, need):
        """Evict idle scenes (least recently used first) until ``need``
        bytes fit; False if they cannot.  (Scenes still uploading are not
        in the store yet, so they are never evicted.)"""
        if self.used + need <= self.budget:
            return True
        for sid in list(self.scenes):
            if self.used + need <= self.budget:
                break
            sc = self.scenes[sid]
            if sc.refs == 0:
                del self.scenes[sid]
                self._account(-self.scene_bytes.pop(sid))
        return self.used + need <= self.budget

    def _build_scene(self, layers):
        """Worker thread: prepare and upload a scene on the upload stream."""
        if not self.cuda:
            return DeviceScene(layers, self.device), None
        with self.gate.work(), torch.cuda.stream(self.upload_stream):
            sc = DeviceScene(layers, self.device)
            ev = torch.cuda.Event()
            ev.record(self.upload_stream)
        return sc, ev

    def _admit(self, req, now):
        """Reserve the request's memory; returns a _Pending (its scene may
        still be uploading) or None when it does not fit yet."""
        sid = req.scene_id
        resident = sid in self.scenes
        uploading = sid in self.uploads
        need = ActiveRequest.estimate(req)
        scene_need = 0 if (resident or uploading) else DeviceScene.estimate(req.workload.layers)
        if need + scene_need > self.budget:
            raise MemoryError(f"request {req.req_id} needs {need + scene_need} B > budget "
                              f"{self.budget} B")
        if resident:
            self.scenes[sid].refs += 1   # pin before making room
        ok = self._make_room(need + scene_need)
        if resident:
            self.scenes[sid].refs -= 1
        if not ok:
            return None
        if resident:
            self.scenes.move_to_end(sid)
            self.scenes[sid].refs += 1
        elif not uploading:
            self._account(scene_need)
            self.uploads[sid] = (self.pool.submit(self._build_scene, req.workload.layers),
                                 scene_need)
        self._account(need)
        return _Pending(req, now, need, uploaded=not resident)

    def _activate(self, p):
        """A pending request whose scene is resident: build its buffers."""
        sid = p.req.scene_id
        if sid not in self.scenes:
            fut, reserved = self.uploads[sid]
            if not fut.done():
                return None
            sc, ev = fut.result()
            del self.uploads[sid]
            if ev is not None:
                torch.cuda.current_stream(self.device).wait_event(ev)
            self.scenes[sid] = sc
            self.scene_bytes[sid] = sc.nbytes()
            self._account(sc.nbytes() - reserved)
            sc.refs += 1
        elif p.uploaded:
            self.scenes[sid].refs += 1   # a later request for a scene uploaded meanwhile
        sc = self.scenes[sid]
        a = ActiveRequest(p.req, sc, self.device, self.precision, p.admitted)
        a.uploaded = p.uploaded
        a.bytes = p.bytes
        return a

    # ----------------------------------------------------------------- steps
    def _step(self, runnable):
        share = self.step_pairs / len(runnable)
        jobs = []
        for a in runnable:
            if len(jobs) + len(a.layer_facets) > self.max_jobs:
                break                       # the rest run in the next step
            jobs += a.take(share, self.tt)
        if self.graphs and self._graph_launch(jobs):
            self.graph_steps += 1
        else:
            self.runner.launch(jobs, self.precision, trace_tile=self.tt, n_blocks=self.n_blocks)
        self.steps += 1
        return jobs

    def _graph_launch(self, jobs):
        """Replay a captured step for ``jobs``; False when they exceed the
        graphs' capacity (the caller then launches eagerly)."""
        _, refracted_items, cls = pack_info(jobs, self.tt)
        if len(jobs) > self.max_jobs or refracted_items > self.max_items:
            return False
        pair = self._graph_sets.get(cls)
        if pair is None:
            pair = self._graph_sets.setdefault(cls, [self._capture(cls), self._capture(cls)])
        sg = pair[self._graph_turn & 1]
        self._graph_turn += 1
        sg.launch(jobs)
        return True

    def _capture(self, cls):
        with self.gate.capture():
            return StepGraph(self.device, self.precision, self.tt, cls, max_jobs=self.max_jobs,
                             max_items=self.max_items, n_blocks=self.n_blocks)

    def _finish_launch(self, a):
        """Queue range compression and the copy to pinned host memory after
        the request's last step (same stream, so it follows that step)."""
        if a.pulse is None:
            a.out.copy_(a.raw)
        else:
            n = a.req.workload.n_samples
            dv.range_compress(a.raw.view(-1, n, 2), a.pulse, out=a.out.view(-1, n, 2))
        if self.cuda:
            a.host_out = torch.empty(a.out.shape, dtype=torch.float64, pin_memory=True)
            a.host_drop = torch.empty(a.dropped.shape, dtype=torch.float64, pin_memory=True)
            a.host_out.copy_(a.out, non_blocking=True)
            a.host_drop.copy_(a.dropped, non_blocking=True)
            a.event = torch.cuda.Event()
            a.event.record()
        else:
            a.host_out = a.out.clone()
            a.host_drop = a.dropped.clone()

    def _done(self, a):
        return a.event is None or a.event.query()

    def _release(self, a, now, results):
        o = a.host_out.numpy()
        results.append(Result(req_id=a.req.req_id, scene_id=a.req.scene_id,
                              field=o[..., 0] + 1j * o[..., 1] if self.keep_fields else None,
                              dropped=a.host_drop.numpy().copy(), pairs=a.pairs,
                              arrival=a.req.arrival if self.realtime else 0.0,
                              admitted=a.admitted, done=now, uploaded=a.uploaded,
                              steps=a.n_steps))
        a.host_out = a.host_drop = None
        a.scene.refs -= 1
        self._account(-a.bytes)
        a.raw = a.out = a.dropped = a.traces = None
        a.keep = []

    # ------------------------------------------------------------------ loop
    def serve(self, requests, realtime=True, clock=time.perf_counter, keep_fields=True):
        """Run a request stream to completion.  ``realtime``: requests arrive
        at their ``arrival`` times on ``clock``; otherwise all are present
        at once (in arrival order, latency counted from the start).
        ``keep_fields``: keep each result's field on the host (else only its
        dropped power and timings).  Returns the results in completion
        order."""
        self.realtime = realtime
        self.keep_fields = keep_fields
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active, finishing, steps_in_flight = [], [], collections.deque()
        results = []
        t0 = clock()

        def now():
            return clock() - t0 if realtime else math.inf

        staged = []
        while pending or waiting or staged or active or finishing:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            while waiting:
                p = self._admit(waiting[0], clock() - t0)
                if p is None:
                    break
                waiting.popleft()
                staged.append(p)
            for p in list(staged):
                a = self._activate(p)
                if a is not None:
                    staged.remove(p)
                    active.append(a)
            runnable = [a for a in active if a.remaining() > 0]
            if runnable:
                while self.cuda and len(steps_in_flight) >= self.max_inflight:
                    steps_in_flight.popleft().synchronize()
                self._step(runnable)
                if self.cuda:
                    ev = torch.cuda.Event()
                    ev.record()
                    steps_in_flight.append(ev)
                for a in runnable:
                    if a.remaining() == 0:
                        self._finish_launch(a)
                        active.remove(a)
                        finishing.append(a)
            progressed = bool(runnable)
            for a in list(finishing):
                if self._done(a):
                    finishing.remove(a)
                    self._release(a, clock() - t0, results)
                    progressed = True
            if not progressed:
                if (finishing and self.cuda) or staged:
                    time.sleep(0.0002)
                elif pending and realtime and not waiting:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.002)))
                elif waiting and not active and not finishing:
                    raise MemoryError("a waiting request can never be admitted")
        return results


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    if not len(lat):
        return {}
    span = max(r.done for r in results) - min(r.arrival for r in results)
    pairs = sum(r.pairs for r in results)
    return {"n": int(len(lat)), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "max_s": float(lat.max()), "pairs": int(pairs),
            "sustained_pairs_per_s": float(pairs / span) if span > 0 else None}


def request_stream(n_requests, lam, seed=7, scale=1.0, n_resident=8, reuse=0.5,
                   n_samples=None, max_facets=None):
    """The serving stream: Poisson arrivals at ``lam`` requests/s, sizes
    drawn as in the uneven batched mix; a fraction ``reuse`` observes one of
    ``n_resident`` scenes (kept resident) with a new instrument and
    trajectory, the rest bring a new scene.  Returns (resident scenes
    {scene_id: layers}, requests)."""
    rng = np.random.default_rng(seed)
    resident = {}
    depths = {}
    for s in range(n_resident):
        layers, depth = draw_scene(rng, (seed, 1000 + s), scale,
                                   max_facets)
        resident[f"resident{s}"] = layers
        depths[f"resident{s}"] = depth
    reqs = []
    t = 0.0
    for i in range(n_requests):
        t += rng.exponential(1.0 / lam)
        if rng.uniform() < reuse:
            sid = f"resident{int(rng.integers(n_resident))}"
            layers, depth = resident[sid], depths[sid]
        else:
            sid = f"new{i}"
            layers, depth = draw_scene(rng, (seed, i), scale, max_facets)
        wl = draw_instrument(rng, sid, layers, depth, scale, n_samples,
                             cross_offset=float(rng.uniform(-200.0, 200.0)))
        reqs.append(Request(req_id=i, scene_id=sid, workload=wl, arrival=t))
    return resident, reqs

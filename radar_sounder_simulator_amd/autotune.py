"""Tile-shape autotuner for the batched persistent launch.

``ss_batch_field`` has two shape parameters: the trace tile (traces per
workgroup: 4, 8 or 16 -- larger tiles reuse each facet's registers over more
traces but leave fewer work items to balance uneven jobs) and the number of
persistent workgroups (2, 3, 4, 6 or 8 per CU: the pair kernels run 1 to 4
waves per SIMD, so a CU holds 1 to 4 of their workgroups at once).  ``tune_batch`` times
every candidate on a proxy of the job list -- the same jobs cut to their
first ``probe_traces`` traces (every facet tile, every job kind; a few
percent of the work of a large mix) -- with HIP events around each launch on
the launch stream, best of ``reps`` after a warm-up, and returns the fastest;
``Tuner`` caches the choice in a JSON file keyed by device, precision and a
coarse signature of the job mix (log2 buckets of total pairs and of the
mean facets per job, layer kinds), so a serving loop tunes each kind of mix
once.  The reference has no counterpart (JAX picks its own launch shapes).
"""

import json
import math
import os

import torch

from . import _lib

TRACE_TILES = (4, 8, 16)
BLOCKS_PER_CU = (2, 3, 4, 6, 8)


def mix_signature(jobs, precision):
    """Coarse key of a job list: precision, log2 bucket of total pairs and of
    mean facets per job, the refracted-job and joint-solver fractions
    (quarters)."""
    n = max(len(jobs), 1)
    pairs = sum(j.n_facets * j.n_traces for j in jobs)
    facets = sum(j.n_facets for j in jobs) / n
    refr = sum(1 for j in jobs if j.kind == 1) / n
    joint = sum(1 for j in jobs if j.kind == 1 and j.refracted.refraction != 0) / n
    return (f"{precision}|p{int(math.log2(max(pairs, 1)))}|f{int(math.log2(max(facets, 1)))}"
            f"|r{int(round(refr * 4))}|j{int(round(joint * 4))}")


def _cus(device):
    if device.type != "cuda":
        return 2
    return torch.cuda.get_device_properties(device).multi_processor_count


def proxy_jobs(jobs, probe_traces=32):
    """Copies of ``jobs`` cut to their first ``probe_traces`` traces (outputs
    to the same rows), the tuning proxy of a large mix."""
    out = []
    for j in jobs:
        c = _lib.Job.from_buffer_copy(j)
        c.n_traces = min(c.n_traces, probe_traces)
        out.append(c)
    return out


def tune_batch(runner, jobs, precision, reps=2, candidates=None, probe_traces=32):
    """Time each (trace_tile, n_blocks) on the proxy of ``jobs`` with
    ``runner`` (a BatchRunner); returns (best (tt, n_blocks), {candidate:
    ms}).  ``probe_traces=None`` times the jobs themselves."""
    dev = runner.device
    cus = _cus(dev)
    if probe_traces:
        jobs = proxy_jobs(jobs, probe_traces)
    if candidates is None:
        candidates = [(tt, k * cus) for tt in TRACE_TILES for k in BLOCKS_PER_CU]
    times = {}
    for tt, nb in candidates:
        runner.launch(jobs, precision, trace_tile=tt, n_blocks=nb)     # warm-up
        best = math.inf
        for _ in range(reps):
            if dev.type == "cuda":
                st = torch.cuda.current_stream(dev)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                runner.launch(jobs, precision, trace_tile=tt, n_blocks=nb)
                e1.record(st)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            else:
                import time
                t0 = time.perf_counter()
                runner.launch(jobs, precision, trace_tile=tt, n_blocks=nb)
                best = min(best, (time.perf_counter() - t0) * 1e3)
        times[(tt, nb)] = best
    return min(times, key=times.get), times


class Tuner:
    """Cached tuning decisions (JSON at ``path``; default
    ``~/.cache/soundersim_hip/tune.json``)."""

    def __init__(self, path=None):
        self.path = path or os.path.join(os.path.expanduser("~"), ".cache", "soundersim_hip",
                                         "tune.json")
        self.table = {}
        if os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    def _key(self, runner, jobs, precision):
        dev = runner.device
        name = torch.cuda.get_device_name(dev) if dev.type == "cuda" else "cpu"
        return f"{name}|{_lib.load().ss_version().decode()}|{mix_signature(jobs, precision)}"

    def batch_shape(self, runner, jobs, precision, reps=2):
        """(trace_tile, n_blocks) for this mix: cached, or tuned now."""
        key = self._key(runner, jobs, precision)
        if key in self.table:
            tt, nb = self.table[key]["best"]
            return int(tt), int(nb)
        best, times = tune_batch(runner, jobs, precision, reps=reps)
        self.table[key] = {"best": list(best),
                           "ms": {f"{tt}x{nb}": ms for (tt, nb), ms in times.items()}}
        self.save()
        return best

    def save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
        except OSError:
            pass

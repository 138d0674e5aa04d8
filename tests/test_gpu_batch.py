"""GPU: the batched persistent launch (ss_batch_field) against one launch
per layer (ss_surface_field / ss_refracted_field) on the same resident
scenes, both tiers, every trace tile, and with the float64 redo forced."""

import numpy as np
import pytest
import torch

from radar_sounder_simulator_amd import _lib, batch, workloads
from radar_sounder_simulator_amd.pipeline import ResidentScene

pytestmark = pytest.mark.gpu


def _scenes():
    wls = [workloads.config2(n=24, n_traces=9, n_samples=4096, seed=3),
           workloads.config1(n=40, n_traces=5, n_samples=512, seed=8),
           workloads.config2(n=17, n_traces=21, n_samples=4096, seed=5)]
    return [ResidentScene(w, device="cuda") for w in wls]


def _single(scenes, precision):
    outs = []
    for rs in scenes:
        for j in range(len(rs.layers)):
            rs.run_layer(j, precision)
        torch.cuda.synchronize()
        outs.append((rs.raw.cpu().numpy().copy(), rs.dropped.cpu().numpy().copy()))
    return outs


def _batched(scenes, precision, tt, redo_all=False):
    br = batch.BatchRunner(device="cuda")
    jobs, keep = [], []
    for rs in scenes:
        j, k = rs.jobs(precision)
        jobs += j
        keep += k
    if redo_all:
        for jb in jobs:
            if jb.kind == 1:
                jb.refracted.flags |= _lib.SS_REDO_ALL
    br.launch(jobs, precision, trace_tile=tt)
    torch.cuda.synchronize()
    return [(rs.raw.cpu().numpy().copy(), rs.dropped.cpu().numpy().copy()) for rs in scenes]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("tt", [4, 8, 16])
def test_batch_fp64_matches_single(tt):
    print(f"case batch/fp64/tt{tt}", flush=True)
    scenes = _scenes()
    ref = _single(scenes, "fp64")
    got = _batched(scenes, "fp64", tt)
    for (o, d), (ro, rd) in zip(got, ref):
        assert all(np.linalg.norm(ro[j]) > 0 for j in range(ro.shape[0]))
        assert _rel(o, ro) <= 1e-12
        np.testing.assert_allclose(d, rd, rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("tt", [4, 8, 16])
def test_batch_mixed_matches_single(tt):
    print(f"case batch/mixed/tt{tt}", flush=True)
    scenes = _scenes()
    ref = _single(scenes, "mixed")
    got = _batched(scenes, "mixed", tt)
    for (o, d), (ro, rd) in zip(got, ref):
        assert _rel(o, ro) <= 1e-5       # same pair arithmetic, FP32 LDS sum order
        np.testing.assert_allclose(d, rd, rtol=1e-5, atol=1e-30)


def test_batch_redo_all_matches_fp64():
    print("case batch/redo_all", flush=True)
    scenes = _scenes()
    ref = _single(scenes, "fp64")
    got = _batched(scenes, "mixed", 8, redo_all=True)
    for (o, d), (ro, rd) in zip(got, ref):
        # refracted layers redone in float64; surface layers stay mixed
        for j in range(o.shape[0]):
            tol = 1e-12 if j > 0 else 1e-4   # surface layer stays on the mixed tier
            assert _rel(o[j], ro[j]) <= tol, (j, _rel(o[j], ro[j]))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_captured_step_replays_match_eager(precision):
    """The HIP-graph-captured step (pinned table -> device table + the step's
    kernels) replayed several times, with eager work and synchronisations in
    between and a different job table each time, equals the eager batched
    launch every time."""
    print(f"case batch/graph/{precision}", flush=True)
    scenes = _scenes()
    tol = 1e-12 if precision == "fp64" else 1e-5
    ref = _batched(scenes, precision, 8)
    jobs, keep = [], []
    for rs in scenes:
        j, k = rs.jobs(precision)
        jobs += j
        keep += k
    sg = batch.StepGraph("cuda", precision, 8, feature_class=2, max_jobs=16,
                         max_items=1 << 14)
    assert sg.graph is not None and sg.fits(jobs)[0]
    for rep in range(3):
        subset = jobs if rep != 1 else jobs[:3]
        for rs in scenes:
            rs.raw.fill_(-1.0)                   # eager work between replays
        torch.cuda.synchronize()
        sg.launch(subset)
        torch.cuda.synchronize()
        done = 0
        for rs, (r, _) in zip(scenes, ref):
            for j in range(rs.raw.shape[0]):
                got = rs.raw[j].cpu().numpy()
                if done < len(subset):
                    assert _rel(got, r[j]) <= tol, (rep, done)
                else:
                    assert np.all(got == -1.0)
                done += 1


@pytest.mark.parametrize("flat", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_joint_classes_match_single(flat, precision):
    """Joint-solver jobs (the general joint solve, and its flat-stack form)
    in the persistent launch against one launch per layer, every trace tile,
    beside a sequential scene in the same batch."""
    print(f"case batch/joint/flat{int(flat)}/{precision}", flush=True)
    wls = [workloads.config3(nx=32, ny=16, n_traces=11, n_firn=4, horizons=(500.0,),
                             flat_stack=flat),
           workloads.config2(n=17, n_traces=21, n_samples=4096, seed=5)]
    scenes = [ResidentScene(w, device="cuda") for w in wls]
    ref = _single(scenes, precision)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for tt in (4, 8, 16):
        got = _batched(scenes, precision, tt)
        for (o, d), (ro, rd) in zip(got, ref):
            assert all(np.linalg.norm(ro[j]) > 0 for j in range(ro.shape[0]))
            assert _rel(o, ro) <= tol, tt
            np.testing.assert_allclose(d, rd, rtol=tol, atol=1e-30)

"""Serving loop host logic on the CPU (kernels through the SIMT emulation of
tests/native/simt): admission by footprint against the budget, eviction of
idle resident scenes, continuous batching in trace chunks, and each result
the same alone and in the mix."""

import numpy as np
import pytest

import simt_harness as sh
from radar_sounder_simulator_amd import _lib
from radar_sounder_simulator_amd import serve as S


@pytest.fixture(scope="module")
def sim():
    return sh.lib()


def _stream(n=6, seed=3):
    return S.request_stream(n, lam=50.0, seed=seed, scale=0.004, n_resident=2, reuse=0.5,
                            n_samples=512, max_facets=150)


def test_request_stream_draws():
    resident, reqs = S.request_stream(40, lam=10.0, seed=1, scale=0.01, n_resident=8,
                                      n_samples=256, max_facets=400)
    assert len(resident) == 8 and len(reqs) == 40
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    reuse = [r for r in reqs if r.scene_id in resident]
    assert 10 <= len(reuse) <= 30
    for r in reuse:                       # a resident scene's layers, new instrument
        assert r.workload.layers is resident[r.scene_id]
    f0 = {r.workload.meta["f0"] for r in reqs}
    assert len(f0) == len(reqs)


def _run(sim, reqs, resident, precision, budget, step_pairs):
    with _lib.use(sim):
        srv = S.Server(device="cpu", precision=precision, budget_bytes=budget,
                       step_pairs=step_pairs, trace_tile=4, n_blocks=2)
        for sid, layers in resident.items():
            srv.add_scene(sid, layers)
        res = srv.serve(reqs, realtime=False)
    return srv, {r.req_id: r for r in res}


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_results_same_alone_and_in_mix(sim, precision):
    resident, reqs = _stream()
    srv, mix = _run(sim, reqs, resident, precision, budget=64 << 20, step_pairs=800)
    assert sorted(mix) == [r.req_id for r in reqs]
    assert max(r.steps for r in mix.values()) > 1     # requests shared several steps
    tol = 1e-12 if precision == "fp64" else 1e-5
    for r in reqs[:4]:
        _, alone = _run(sim, [r], resident, precision, budget=64 << 20, step_pairs=1e12)
        a, m = alone[r.req_id], mix[r.req_id]
        assert np.linalg.norm(a.field) > 0
        assert np.linalg.norm(m.field - a.field) <= tol * np.linalg.norm(a.field)
        np.testing.assert_allclose(m.dropped, a.dropped, rtol=tol, atol=1e-30)


def test_admission_respects_budget(sim):
    resident, reqs = _stream(n=5, seed=9)
    need = [S.ActiveRequest.estimate(r) + S.DeviceScene.estimate(r.workload.layers)
            for r in reqs]
    budget = int(max(need) * 1.2)          # at most about one request at a time
    srv, res = _run(sim, reqs, resident, "mixed", budget=budget, step_pairs=5000)
    assert len(res) == len(reqs)
    assert srv.peak <= budget
    assert all(r.uploaded == (r.scene_id not in resident) or r.scene_id in resident
               for r in res.values())
    with pytest.raises(MemoryError):
        _run(sim, reqs[:1], {}, "mixed", budget=min(need) // 2, step_pairs=5000)


def test_captured_step_capacity_is_bounded(sim):
    """A huge step share (one request per step) must not size the captured
    steps' redo workspace from the share: it is capped, and larger steps run
    as eager launches."""
    with _lib.use(sim):
        srv = S.Server(device="cpu", precision="mixed", budget_bytes=64 << 20,
                       step_pairs=1e12, trace_tile=8, n_blocks=2)
        assert srv.max_items == 1 << 22
        ws = _lib.load().ss_batch_step_workspace_bytes(_lib.SS_PREC_MIXED, srv.max_items, 8)
        assert ws <= 1100 << 20


def test_capture_gate_excludes_workers():
    """A capture waits for workers inside ``work()`` and holds new ones off
    until it ends; workers run together otherwise."""
    import threading
    import time

    gate = S.CaptureGate()
    log, lock = [], threading.Lock()

    def worker(tag, hold):
        with gate.work():
            with lock:
                log.append(("in", tag))
            time.sleep(hold)
            with lock:
                log.append(("out", tag))

    a = threading.Thread(target=worker, args=("a", 0.3))
    b = threading.Thread(target=worker, args=("b", 0.3))
    a.start(); b.start()
    time.sleep(0.1)
    with gate.capture():
        with lock:
            assert ("out", "a") in log and ("out", "b") in log   # both left first
            log.append(("capture", None))
        c = threading.Thread(target=worker, args=("c", 0.0))
        c.start()
        time.sleep(0.2)
        with lock:
            assert ("in", "c") not in log                        # held off meanwhile
    c.join(); a.join(); b.join()
    assert log.index(("in", "c")) > log.index(("capture", None))
    assert log.index(("in", "b")) < log.index(("out", "a"))      # a and b overlapped

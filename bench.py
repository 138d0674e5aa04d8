"""Benchmark: facet.trace evaluations/s of the coherent facet kernel on one MI355X.

Workload: the surface + bed refraction scene (512 x 512 facets per interface
at 25 m, 256 traces at 20 km, 4096 samples at 48 MHz, exponential sub-facet
roughness + grazing fix, cos^2 pattern, 60 MHz / 15 MHz chirp; see
radar_sounder_simulator_amd/workloads.py config2), synthetic seeded data,
resident in HBM before the timed region.  One step = one full pass of the
hot path: surface kernel + refracted bed kernel + range compression of both
layers.  The headline value is the mixed-precision tier; the plain float64
tier is timed beside it on the same scene, and the mixed result's error
against it (max dB / max mrad over bins within 30 dB of each trace's peak)
is reported.  The float64 tier reproduces the float64 oracle to rel L2
<= 1e-9 (tests/test_gpu_parity.py), so it is the stand-in for the oracle at
this size.

Output: one JSON line (rank 0).  ``roofline`` covers the dominant kernel
(the refracted bed kernel) with its duration measured by HIP events on the
launch stream; ``cpu_baseline`` times this project's float64 C restatement
(oracle/csrc/oracle.c, OpenMP over traces) on a bounded sample of traces.

Two further sections ride on the same line:
  ``batched``  the uneven 64-scene mix (facets log-uniform in [1e4, 2e6],
               traces uniform in [50, 2000], 1-4 layers, 30-330 MHz), sizes
               times ``--batch-scale``, in one persistent launch with the
               autotuned tile/grid, against the same scenes run one launch
               per layer, scene after scene;
  ``serving``  a Poisson request stream (``--lam`` requests/s, sizes as the
               batched mix times ``--serve-scale``, half of the requests on 8
               resident scenes with a new instrument and trajectory), served by
               the continuous-batching loop on the mixed tier and on the
               float64 tier: p50/p99 latency (arrival to result in pinned
               host memory) and sustained pairs/s;
  ``multilayer`` the 25-interface firn/horizon/bed scene (workloads.config3)
               over ``--ml-traces`` of its traces: both tiers, per-layer
               kernel times and the mixed tier's error per layer;
  ``incoherent`` the power-summing kernel on the surface layer of the same
               scene: both tiers, the mixed tier's error, roofline, and the
               NumPy oracle timed on 8 traces;
  ``processing`` focused SAR (time-domain backprojection, 2 km hann
               aperture) of the scene's compressed surface field, both tiers,
               presum / unfocused boxcars, and the NumPy oracle on the
               first 41 traces.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def errors_db_mrad(mixed, ref, floor_db=30.0):
    """Max power error (dB) and phase error (mrad) over bins within
    ``floor_db`` of each (layer, trace)'s peak."""
    a = mixed[..., 0] + 1j * mixed[..., 1]
    b = ref[..., 0] + 1j * ref[..., 1]
    pk = np.abs(b).max(axis=-1, keepdims=True)
    m = np.abs(b) > pk * 10 ** (-floor_db / 20.0)
    db = np.abs(20 * np.log10(np.abs(a[m]) / np.abs(b[m]))).max()
    mrad = np.abs(np.angle(a[m] / b[m])).max() * 1e3
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    return float(db), float(mrad), float(rel)


def cpu_baseline(wl, n_traces):
    """Time the C float64 oracle on ``n_traces`` traces x every facet of both
    layers, threads = available host cores (capped at the trace count)."""
    from oracle import cbuild, cref
    if not os.path.exists(cbuild.OUT):
        cbuild.build()
    cores = len(os.sched_getaffinity(0))
    nth = max(1, min(cores, n_traces))
    tr = wl.track
    idx = np.linspace(0, len(tr.positions) - 1, n_traces).round().astype(int)
    pos = tr.positions[idx]
    pat = None
    if wl.pattern is not None:
        k, pv, pa, pb = wl.pattern
        pat = (k, pv[idx], pa, pb)
    s, b = wl.layers
    t0 = time.perf_counter()
    f = s.facets
    cref.coherent_cluttergram(pos, tr.u_ct[idx], f.centers, f.normals, f.areas, f.e1, f.e2,
                              k=wl.k0, gamma=s.gamma, t0=wl.t0, dt=wl.dt,
                              n_samples=wl.n_samples, c=wl.c, pattern=pat,
                              roughness=s.roughness, taper_s=wl.taper_s,
                              d_phi_area=wl.d_phi_area, nthreads=nth)
    cref.refracted_cluttergram(pos, tr.u_ct[idx], b.facets, b.crossed, b.eps_leg, b.att_leg,
                               t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c,
                               gamma=b.gamma, k0=wl.k0, pattern=pat, roughness=b.roughness,
                               taper_s=wl.taper_s, d_phi_area=wl.d_phi_area, nthreads=nth)
    el = time.perf_counter() - t0
    pairs = n_traces * (len(s.facets) + len(b.facets))
    cpu = open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(": ")
    return {"value": pairs / el, "unit": "facet*trace/s", "cores": nth, "kind": "port",
            "sample": f"{n_traces} of {len(tr.positions)} traces x all "
                      f"{len(s.facets)}+{len(b.facets)} facets (surface+bed), float64 C "
                      f"restatement, OpenMP {nth} threads on {cpu}; {el:.1f} s"}


def bench_batched(scale, dev, seed=4, fp64_batched=False):
    """One persistent launch over the 64-scene mix vs one launch per layer, on
    the mixed tier; the float64 tier as one plain launch per layer (the
    comparison the batched mix is quoted on), and its persistent launch only
    when ``fp64_batched``."""
    import torch
    from radar_sounder_simulator_amd import autotune
    from radar_sounder_simulator_amd.batch import BatchRunner, uneven_workloads
    from radar_sounder_simulator_amd.pipeline import ResidentScene
    t = time.time()
    wls = uneven_workloads(64, seed=seed, scale=scale)
    log(f"bench: batched mix drawn in {time.time() - t:.1f} s")
    scenes = []
    for i, w in enumerate(wls):
        scenes.append(ResidentScene(w, device=dev))
        if i % 16 == 15:
            log(f"bench: batched mix: {i + 1} scenes resident ({time.time() - t:.1f} s)")
    pairs = sum(sum(rs.pairs_per_layer()) for rs in scenes)
    st = torch.cuda.current_stream()
    out = {"scenes": 64, "scale": scale, "pairs": pairs,
           "layers": sum(len(w.layers) for w in wls)}
    runner = BatchRunner(device=dev)
    tuner = autotune.Tuner()
    for prec in ("mixed", "fp64"):
        jobs, keep = [], []
        for rs in scenes:
            j, k = rs.jobs(prec)
            jobs += j
            keep += k
        t = time.time()
        if prec == "mixed" or fp64_batched:
            tt, nb = tuner.batch_shape(runner, jobs, prec)
            log(f"bench: batched {prec}: tile {tt}, {nb} blocks (tuned in {time.time() - t:.1f} s)")
        else:
            tt = nb = None

        def batched():
            runner.launch(jobs, prec, trace_tile=tt, n_blocks=nb)
            for rs in scenes:
                rs.compress()

        def sequential():
            for rs in scenes:
                rs.run(prec)

        res = {"trace_tile": tt, "n_blocks": nb}
        runs = (("batched", batched), ("sequential", sequential))
        if prec == "fp64" and not fp64_batched:
            runs = runs[1:]
        for name, fn in runs:
            fn()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            fn()
            e1.record(st)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            res[name] = {"ms": ms, "pairs_per_s": pairs / (ms * 1e-3)}
            log(f"bench: batched {prec} {name}: {ms:.1f} ms")
        if "batched" in res:
            res["speedup"] = res["sequential"]["ms"] / res["batched"]["ms"]
            log(f"bench: batched {prec}: {res['batched']['ms']:.1f} ms vs sequential "
                f"{res['sequential']['ms']:.1f} ms (tile {tt}, {nb} blocks)")
        out[prec] = res
        torch.cuda.empty_cache()
    # the comparison the batched configuration names: the mixed persistent
    # launch against the same scenes as one plain float64 launch per layer
    out["mixed_batched_vs_fp64_sequential"] = (out["fp64"]["sequential"]["ms"] /
                                               out["mixed"]["batched"]["ms"])
    out["joint_layers"] = sum(1 for w in wls for L in w.layers if L.refraction == "joint")
    log(f"bench: batched mixed vs sequential fp64: {out['mixed_batched_vs_fp64_sequential']:.2f}x")
    return out


def errors_db_mrad_dev(mixed, ref, floor_db=30.0):
    """errors_db_mrad on device tensors (..., n, 2) float64."""
    import torch
    a = torch.view_as_complex(mixed.contiguous())
    b = torch.view_as_complex(ref.contiguous())
    ab = b.abs()
    pk = ab.amax(dim=-1, keepdim=True)
    m = ab > pk * 10 ** (-floor_db / 20.0)
    r = a[m] / b[m]
    db = (20 * torch.log10(r.abs())).abs().max().item() if r.numel() else 0.0
    mrad = r.angle().abs().max().item() * 1e3 if r.numel() else 0.0
    rel = ((a - b).abs().pow(2).sum().sqrt() / ab.pow(2).sum().sqrt()).item()
    return float(db), float(mrad), float(rel)


def bench_multilayer(n_traces, dev, sub_traces=8):
    """The multilayer scene (2048 x 1024 facets per interface, 25 target
    interfaces: rough surface, 20 firn sublayers, 3 horizons, rough bed;
    8192 samples, 150/30 MHz Tukey chirp) over ``n_traces`` of its 1000
    traces, the joint layers in their flat-stack form (the opt-in the scene
    is built for): one mixed and one float64 pass, per-layer kernel times and
    the mixed tier's error per layer against the float64 tier.  Beside it,
    the reference's general joint solve on the bed (24 crossed interfaces)
    and a mid-stack firn layer over ``sub_traces`` traces and all facets:
    its rate in both tiers, and how many lanes its validity masks and bins
    differ from the flat-stack form's (float64 tier of both)."""
    import torch
    from radar_sounder_simulator_amd import device as dv
    from radar_sounder_simulator_amd import workloads
    from radar_sounder_simulator_amd.pipeline import ResidentScene
    t = time.time()
    wl = workloads.config3(n_traces=n_traces)
    rs = ResidentScene(wl, device=dev)
    log(f"bench: multilayer scene built and uploaded in {time.time() - t:.1f} s")
    pairs_l = rs.pairs_per_layer()
    pairs = sum(pairs_l)
    st = torch.cuda.current_stream()
    names = [L.name for L in wl.layers]
    out = {"workload": "multilayer, 2048x1024 facets at 30 m x 25 interfaces, "
                       f"{n_traces} of 1000 traces at 20 km, 8192 samples, 150/30 MHz, "
                       "joint layers in the flat-stack form",
           "traces": n_traces, "layers": len(names), "pairs_per_step": pairs}
    kept = None
    for prec in ("mixed", "fp64"):
        for j in range(3):             # load each kernel once before timing
            rs.run_layer(j, prec)
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 2)]
        e[0].record(st)
        for j in range(len(names)):
            rs.run_layer(j, prec)
            e[j + 1].record(st)
        rs.compress()
        e[-1].record(st)
        torch.cuda.synchronize()
        ms = e[0].elapsed_time(e[-1])
        per = [e[j].elapsed_time(e[j + 1]) for j in range(len(names))]
        out[prec] = {"ms_per_step": ms, "pairs_per_s": pairs / (ms * 1e-3),
                     "layer_ms": dict(zip(names, per))}
        if prec == "mixed":
            kept = rs.raw.clone()
        log(f"bench: multilayer {prec}: {ms:.1f} ms/step, {pairs / (ms * 1e-3):.3e} pairs/s")
    errs = {}
    for j, n in enumerate(names):
        db, mrad, rel = errors_db_mrad_dev(kept[j], rs.raw[j])
        errs[n] = {"max_db": db, "max_mrad": mrad, "rel_l2": rel}
    del kept
    out["mixed_vs_fp64_per_layer"] = errs
    out["max_db"] = max(e["max_db"] for e in errs.values())
    out["max_mrad"] = max(e["max_mrad"] for e in errs.values())
    out["speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    log(f"bench: multilayer mixed vs fp64 worst layer {out['max_db']:.2e} dB "
        f"{out['max_mrad']:.3f} mrad, speedup {out['speedup_vs_fp64']:.2f}x")
    # the reference's general joint solve on a trace subset, all facets
    tr = wl.track
    idx = np.linspace(0, n_traces - 1, sub_traces).round().astype(int)
    trs = dv.DeviceTraces(tr.positions[idx], tr.u_ct[idx], None, wl.pattern[1][idx],
                          device=dev)
    gen = {"traces": sub_traces, "layers": {}}
    for name in ("firn5", "firn12", "bed"):
        j = names.index(name)
        L, fac, ifc, rough = rs.layers[j]
        kw = dict(mode="coherent", t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c,
                  gamma=L.gamma, k0=wl.k0, pattern=wl.pattern, refraction="joint")
        res, bins = {}, {}
        for form, flat in (("general", False), ("flat", True)):
            for prec in ("fp64", "mixed"):
                dv.refracted_field(fac, trs, ifc, L.eps_leg, L.att_leg, precision=prec,
                                   flat_stack=flat, **kw)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                r = dv.refracted_field(fac, trs, ifc, L.eps_leg, L.att_leg, precision=prec,
                                       flat_stack=flat, pair_bins=prec == "fp64", **kw)
                e1.record(st)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1)
                res[(form, prec)] = {"kernel_ms": ms, "pairs_per_s": fac.n * sub_traces / (ms * 1e-3)}
                if prec == "fp64":
                    bins[form] = r[2]
        bg, bf = bins["general"], bins["flat"]
        inv = -2147483648
        vg, vf = bg != inv, bf != inv
        rec = {"crossed": len(L.crossed), "pad_to": L.pad_to, "pairs": fac.n * sub_traces,
               "general": {p: res[("general", p)] for p in ("fp64", "mixed")},
               "flat": {p: res[("flat", p)] for p in ("fp64", "mixed")},
               "validity_differs": int((vg != vf).sum().item()),
               "valid_general_only": int((vg & ~vf).sum().item()),
               "valid_flat_only": int((vf & ~vg).sum().item()),
               "bins_differ_both_valid": int(((bg != bf) & vg & vf).sum().item()),
               "invalid_both": int((~vg & ~vf).sum().item())}
        rec["general_mixed_speedup_vs_fp64"] = (res[("general", "fp64")]["kernel_ms"] /
                                                res[("general", "mixed")]["kernel_ms"])
        gen["layers"][name] = rec
        log(f"bench: multilayer {name} general joint: fp64 {res[('general', 'fp64')]['kernel_ms']:.1f} "
            f"ms, mixed {res[('general', 'mixed')]['kernel_ms']:.1f} ms over {sub_traces} traces; "
            f"validity differs on {rec['validity_differs']} of {rec['pairs']} lanes")
    out["reference_joint_form"] = gen
    del rs
    return out


def bench_incoherent(wl, dev, steps=10, cpu_traces=8):
    """The incoherent (power-summing) kernel on the surface layer of the
    surface+bed scene (512 x 512 facets, 256 traces, 4096 samples, cos^2
    pattern): both tiers, HIP events on the launch stream, the mixed tier's
    error against the float64 tier, and the oracle timed on a few traces."""
    import torch
    from radar_sounder_simulator_amd import device as dv
    from radar_sounder_simulator_amd import roofline
    s = wl.layers[0].facets
    tr = wl.track
    fac = dv.DeviceFacets(s.centers, s.normals, s.areas, s.e1, s.e2, device=dev)
    trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, wl.pattern[1], device=dev)
    kw = dict(t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c, pattern=wl.pattern)
    pairs = fac.n * trs.n
    st = torch.cuda.current_stream()
    out = {"workload": "incoherent surface, 512x512 facets at 25 m, 256 traces at 20 km, "
                       "4096 samples, cos^2 pattern", "pairs_per_step": pairs}
    res = {}
    for prec in ("mixed", "fp64"):
        dv.incoherent_power(fac, trs, precision=prec, **kw)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(st)
        for i in range(steps):
            p, d, _ = dv.incoherent_power(fac, trs, precision=prec, **kw)
            ev[i + 1].record(st)
        torch.cuda.synchronize()
        ms = float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]))
        res[prec] = p.cpu().numpy()
        out[prec] = {"kernel_ms": ms, "pairs_per_s": pairs / (ms * 1e-3)}
        log(f"bench: incoherent {prec}: {ms:.3f} ms, {pairs / (ms * 1e-3):.3e} pairs/s")
    ref = res["fp64"]
    out["mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(res["mixed"] - ref) / np.linalg.norm(ref))
    f = roofline.incoherent_pair_flops("tabulated")
    ach = f * pairs / (out["mixed"]["kernel_ms"] * 1e-3) / 1e12
    kname = "incoherent_kernel<1, 5>"
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(kname, {}).get("bytes")
    out["roofline"] = {"bound": "valu", "flops_per_pair": f, "achieved": ach,
                       "peak": roofline.PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                       "frac": ach / roofline.PEAK_FP32_TFLOPS, "traffic": traffic,
                       "traffic_source": "profiles/incoherent_kernel_surface_512x512_256traces.md",
                       "kernel": kname}
    out["speedup_vs_fp64"] = out["fp64"]["kernel_ms"] / out["mixed"]["kernel_ms"]
    from oracle import incoherent as oi
    t = time.time()
    oi.incoherent_cluttergram(tr.positions[:cpu_traces], tr.u_ct[:cpu_traces], s.centers,
                              s.normals, s.areas, t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples,
                              c=wl.c, pattern=(wl.pattern[0], wl.pattern[1][:cpu_traces],
                                               wl.pattern[2], wl.pattern[3]))
    el = time.time() - t
    out["cpu_baseline"] = {"value": cpu_traces * fac.n / el, "unit": "facet*trace/s",
                           "cores": 1, "kind": "port",
                           "sample": f"oracle/incoherent.py (NumPy float64) over {cpu_traces} "
                                     "traces of the same scene"}
    log(f"bench: incoherent mixed vs fp64 rel {out['mixed_vs_fp64_rel_l2']:.2e}; "
        f"cpu {cpu_traces * fac.n / el:.3e} pairs/s")
    return out


def bench_processing(field, wl, dev, aperture_m=2000.0, steps=5, cpu_traces=41):
    """Focused SAR (backprojection) over the compressed surface field of the
    surface+bed scene (256 traces at 25 m, 4096 bins, 60 MHz) with a 2 km
    hann aperture (81 traces), both tiers; presum (8) and the unfocused
    boxcar over the same 81 traces; the oracle's backprojection timed on the
    first ``cpu_traces`` traces."""
    import torch
    from radar_sounder_simulator_amd import device as dv
    from radar_sounder_simulator_amd import roofline
    T, nb = field.shape[0], field.shape[1]
    pos = wl.track.positions
    seg = np.linalg.norm(np.diff(pos, axis=0), axis=1)
    s = np.concatenate([[0.0], np.cumsum(seg)])
    twtt = wl.t0 + np.arange(nb) * wl.dt
    k = wl.k0
    ft = torch.as_tensor(field, device=dev)
    tw = torch.as_tensor(twtt, device=dev)
    st_ = torch.as_tensor(s, device=dev)
    n_ap = np.array([int(np.sum(np.abs(s - si) <= aperture_m / 2)) for si in s])
    terms = int(n_ap.sum()) * nb
    st = torch.cuda.current_stream()
    out = {"workload": f"focused SAR of the surface field, {T} traces at 25 m, {nb} bins, "
                       f"{aperture_m:.0f} m hann aperture", "terms_per_launch": terms}
    res = {}
    for prec in ("mixed", "fp64"):
        dv.focused_sar(ft, tw, st_, c=wl.c, k=k, aperture_m=aperture_m, precision=prec)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(st)
        for i in range(steps):
            y = dv.focused_sar(ft, tw, st_, c=wl.c, k=k, aperture_m=aperture_m, precision=prec)
            ev[i + 1].record(st)
        torch.cuda.synchronize()
        ms = float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]))
        res[prec] = y.cpu().numpy()
        out[prec] = {"kernel_ms": ms, "terms_per_s": terms / (ms * 1e-3)}
        log(f"bench: focused SAR {prec}: {ms:.3f} ms, {terms / (ms * 1e-3):.3e} terms/s")
    a = res["mixed"][..., 0] + 1j * res["mixed"][..., 1]
    b = res["fp64"][..., 0] + 1j * res["fp64"][..., 1]
    out["mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    f = roofline.FOCUS_TERM_FLOPS
    ach = f * terms / (out["fp64"]["kernel_ms"] * 1e-3) / 1e12
    out["roofline"] = {"bound": "valu", "flops_per_term": f, "achieved": ach,
                       "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                       "frac": ach / roofline.PEAK_FP64_TFLOPS, "kernel": "focus_kernel<0>"}
    for name, n, stride in (("presum_8", 8, 8), ("unfocused_81", 81, 1)):
        n_out = T // n if stride == n else T - n + 1
        dv.slow_time_boxcar(ft, n, stride, n_out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(steps):
            dv.slow_time_boxcar(ft, n, stride, n_out)
        e1.record(st)
        torch.cuda.synchronize()
        out[name] = {"kernel_ms": e0.elapsed_time(e1) / steps,
                     "bytes_read": n_out * n * nb * 16, "bytes_written": n_out * nb * 16}
    from oracle import processing as opr
    m = cpu_traces
    s_sl = s[:m]
    sl_terms = int(sum(np.sum(np.abs(s_sl - si) <= aperture_m / 2) for si in s_sl)) * nb
    t = time.time()
    opr.focused_sar(a[:m].astype(np.complex128), s_sl, twtt, c=wl.c, k=k, aperture_m=aperture_m)
    el = time.time() - t
    out["cpu_baseline"] = {"value": sl_terms / el, "unit": "terms/s", "cores": 1,
                           "kind": "port",
                           "sample": f"oracle/processing.py focused_sar (NumPy float64) on the "
                                     f"first {m} traces ({sl_terms} terms)"}
    log(f"bench: focused SAR mixed vs fp64 rel {out['mixed_vs_fp64_rel_l2']:.2e}; "
        f"cpu {out['cpu_baseline']['value']:.3e} terms/s")
    return out


def bench_serving(n_requests, lam, scale, dev, seed=7, fp64_requests=0):
    """The request stream at ``lam`` requests/s on the mixed tier, and its
    first ``fp64_requests`` requests on the float64 tier (none when 0)."""
    import torch
    from radar_sounder_simulator_amd import serve as S
    t = time.time()
    resident, reqs = S.request_stream(n_requests, lam=lam, seed=seed, scale=scale,
                                      n_resident=8, reuse=0.5)
    log(f"bench: serving stream drawn in {time.time() - t:.1f} s")
    out = {"requests": n_requests, "lambda_per_s": lam, "scale": scale,
           "resident_scenes": 8, "reuse": 0.5,
           "pairs": int(sum(r.workload.pairs() for r in reqs))}
    runs = [("mixed", reqs)] + ([("fp64", reqs[:fp64_requests])] if fp64_requests else [])
    for prec, rq in runs:
        srv = S.Server(device=dev, precision=prec, step_pairs=4e8)
        for sid, layers in resident.items():
            srv.add_scene(sid, layers)
        res = srv.serve(rq, realtime=True, keep_fields=False)
        torch.cuda.synchronize()
        st = S.latency_stats(res)
        st.update({"requests": len(rq), "steps": srv.steps, "graph_steps": srv.graph_steps,
                   "peak_hbm_bytes": srv.peak, "budget_bytes": srv.budget})
        out[prec] = st
        log(f"bench: serving {prec}: p50 {st['p50_s'] * 1e3:.1f} ms p99 "
            f"{st['p99_s'] * 1e3:.1f} ms sustained {st['sustained_pairs_per_s']:.3e} pairs/s")
        del srv
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="facets per side per interface")
    ap.add_argument("--traces", type=int, default=256)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-traces", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--batch-scale", type=float, default=1.0)
    ap.add_argument("--serve-scale", type=float, default=1.0)
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--lam", default="2,20", help="arrival rates (requests/s), comma-separated")
    ap.add_argument("--serve-fp64-requests", type=int, default=32,
                    help="requests of the highest-rate stream also served on the float64 tier")
    ap.add_argument("--fp64-batched", action="store_true",
                    help="also time the float64 tier's persistent launch over the mix")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--ml-traces", type=int, default=1000,
                    help="traces of the multilayer scene's 1000 in its section")
    ap.add_argument("--no-multilayer", action="store_true")
    ap.add_argument("--no-incoherent", action="store_true")
    ap.add_argument("--no-processing", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the coherent facet kernel shards by "
              "traces with no exchange step, and this build measures one MI355X "
              "(--gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from radar_sounder_simulator_amd import roofline, workloads
    from radar_sounder_simulator_amd.pipeline import ResidentScene

    log(f"bench: building workload n={args.n} traces={args.traces} samples={args.samples}")
    wl = workloads.config2(n=args.n, n_traces=args.traces, n_samples=args.samples)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rs = ResidentScene(wl, device=dev)
    pairs_l = rs.pairs_per_layer()
    pairs = sum(pairs_l)
    st = torch.cuda.current_stream()

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            rs.run(precision)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(st)
        for s in range(steps):
            e = evs[s]
            e[0].record(st)
            rs.run_layer(0, precision)
            e[1].record(st)
            rs.run_layer(1, precision)
            e[2].record(st)
            rs.compress()
            e[3].record(st)
        t1.record(st)
        torch.cuda.synchronize()
        total = t0.elapsed_time(t1) / steps
        k_surf = np.mean([e[0].elapsed_time(e[1]) for e in evs])
        k_bed = np.mean([e[1].elapsed_time(e[2]) for e in evs])
        k_cmp = np.mean([e[2].elapsed_time(e[3]) for e in evs])
        return total, k_surf, k_bed, k_cmp

    log("bench: mixed precision")
    ms, ms_s, ms_b, ms_c = timed("mixed", args.steps, args.warmup)
    out_mixed = rs.out.cpu().numpy()
    log(f"bench: mixed {ms:.2f} ms/step (surface {ms_s:.2f}, bed {ms_b:.2f}, "
        f"compress {ms_c:.2f})")
    extra = {}
    if not args.no_fp64:
        log("bench: float64")
        md, md_s, md_b, md_c = timed("fp64", args.fp64_steps, 1)
        out_fp64 = rs.out.cpu().numpy()
        db, mrad, rel = errors_db_mrad(out_mixed, out_fp64)
        per_layer = [errors_db_mrad(out_mixed[j:j + 1], out_fp64[j:j + 1])
                     for j in range(len(pairs_l))]
        log(f"bench: fp64 {md:.2f} ms/step (surface {md_s:.2f}, bed {md_b:.2f}); "
            f"mixed vs fp64 max {db:.2e} dB {mrad:.3f} mrad rel {rel:.2e}")
        extra = {
            "fp64_value": pairs / (md * 1e-3), "fp64_ms_per_step": md,
            "speedup_vs_fp64": md / ms,
            "fp64_kernel_ms": {"surface": md_s, "bed": md_b, "compress": md_c},
            "mixed_vs_fp64": {"max_db": db, "max_mrad": mrad, "rel_l2": rel,
                              "domain": "bins within 30 dB of each trace's peak",
                              "per_layer": {n: {"max_db": e[0], "max_mrad": e[1]}
                                            for n, e in zip(["surface", "bed"], per_layer)}},
        }
    s, b = wl.layers
    f_bed = roofline.refracted_pair_flops(1, "tabulated", True, b.roughness[3])
    f_surf = roofline.surface_pair_flops("tabulated", True, s.roughness[3])
    achieved = f_bed * pairs_l[1] / (ms_b * 1e-3) / 1e12
    # mixed, sequential, area-only D_Phi, tabulated, plain flag profile, 4 waves per SIMD
    kname = "refracted_kernel<1, 0, 1, 5, 1, 4>"
    traffic = None
    prof = {}
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f).get(kname, {})
            traffic = prof.get("bytes")
    roof = {"bound": "valu", "achieved": achieved, "peak": roofline.PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / roofline.PEAK_FP32_TFLOPS,
            "traffic": traffic, "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE "
            "(calibrated by a 1 GiB copy), profiles/refracted_kernel_mixed_traffic_surface_bed.md",
            # algorithmic flops count a transcendental as one and leave out the
            # integer, compare, select and conversion instructions; the VALU
            # issue counters of the same launch show how busy the vector ALUs are
            "valu_busy": prof.get("valu_busy"), "valu_utilization": prof.get("valu_utilization"),
            "valu_source": prof.get("valu_source"),
            # the FP32 peak counts packed FP32 (two per lane per cycle); the
            # kernels issue no packed FP32, whose own issue rate is half of it
            "peak_nonpacked_fp32": roofline.PEAK_FP32_TFLOPS / 2,
            "frac_nonpacked_fp32": achieved / (roofline.PEAK_FP32_TFLOPS / 2),
            # bytes the launch cannot avoid: facets, crossed-surface table,
            # trace records and the field, each once
            "unique_bytes": roofline.refracted_unique_bytes(pairs_l[1] // rs.traces.n,
                                                            rs.traces.n, wl.n_samples),
            "kernel": kname,
            "flops_per_pair": f_bed, "kernel_ms": ms_b,
            "surface_kernel": {"flops_per_pair": f_surf, "kernel_ms": ms_s,
                               "achieved": f_surf * pairs_l[0] / (ms_s * 1e-3) / 1e12}}
    cpu = None
    if not args.no_cpu:
        log(f"bench: cpu baseline ({args.cpu_traces} traces)")
        cpu = cpu_baseline(wl, args.cpu_traces)
    line = {
        "metric": "facet*trace evals/s (mixed precision, surface+bed refraction scene)",
        "value": pairs / (ms * 1e-3), "unit": "facet*trace/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": "surface+bed refraction, 512x512 facets/interface at 25 m, "
                               "256 traces at 20 km, 4096 samples, 60/15 MHz",
                   "facets_per_interface": args.n * args.n, "traces": args.traces,
                   "samples": args.samples, "pairs_per_step": pairs},
        "kernel_ms": {"surface": ms_s, "bed": ms_b, "compress": ms_c},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    if not args.no_incoherent:
        log("bench: incoherent kernel on the surface layer")
        line["incoherent"] = bench_incoherent(wl, dev)
    if not args.no_processing:
        log("bench: focused SAR of the surface field")
        line["processing"] = bench_processing(out_mixed[0], wl, dev)
    del rs
    torch.cuda.empty_cache()
    if not args.no_batch:
        log(f"bench: batched uneven mix (scale {args.batch_scale})")
        line["batched"] = bench_batched(args.batch_scale, dev, fp64_batched=args.fp64_batched)
        torch.cuda.empty_cache()
    if not args.no_serve:
        line["serving"] = {}
        lams = [float(x) for x in str(args.lam).split(",")]
        for lam in lams:
            n64 = args.serve_fp64_requests if lam == max(lams) else 0
            log(f"bench: serving stream ({args.serve_requests} requests at {lam}/s, "
                f"{n64} of them also on the float64 tier)")
            line["serving"][f"lambda_{lam:g}"] = bench_serving(args.serve_requests, lam,
                                                               args.serve_scale, dev,
                                                               fp64_requests=n64)
            torch.cuda.empty_cache()
    if not args.no_multilayer:
        torch.cuda.empty_cache()
        log(f"bench: multilayer scene ({args.ml_traces} traces)")
        line["multilayer"] = bench_multilayer(args.ml_traces, dev)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

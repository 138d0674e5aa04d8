"""The gfx950 kernel source executed on the CPU through the SIMT emulation
(tests/native/simt/hip/hip_runtime.h): one host thread per lane, real
barriers, wavefront shuffles/ballots and atomics.  This checks the kernels'
own bookkeeping -- trace tiles, LDS bin windows and flushes, the deferred
float64 redo (item flags and lane masks), batched persistent launches and
the float64 fold-back -- against the oracle and against each other without
a GPU.  Float32 hardware primitives run as libm here, so mixed-tier values
are checked to the mixed tolerance only.  Sizes are tiny."""

import numpy as np
import pytest
import torch

import simt_harness as sh
from oracle import coherent as oc
from oracle import multilayer as om
from radar_sounder_simulator_amd import _lib
from radar_sounder_simulator_amd import batch as B
from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd import workloads as W
from radar_sounder_simulator_amd.pipeline import ResidentScene
from radar_sounder_simulator_amd.roughness import speckle_phasors
from radar_sounder_simulator_amd.scene import gaussian_surface, grid_facets, straight_track

C = 299792458.0


@pytest.fixture(scope="module")
def sim():
    return sh.lib()


def _bed(seed=5, n=20):
    rng = np.random.default_rng(seed)
    dx = 25.0
    surf = grid_facets(3.0 * gaussian_surface(n + 1, dx, 100.0, rng), dx)
    bed = grid_facets(-300.0 + 30.0 * gaussian_surface(n + 1, dx, 200.0, rng), dx)
    tr = straight_track(5, 40.0, 3000.0)
    return surf, bed, tr


def _refr(sim, precision, extra=0, rough=False):
    surf, bed, tr = _bed()
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    ph = None
    if rough:
        ph = speckle_phasors(len(bed), 3)
        kw.update(roughness=(0.5, 20.0, 24, "exponential"), d_phi_area=True, taper_s=0.05)
    with _lib.use(sim):
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2,
                              phasors=ph, device="cpu")
        ifc = dv.DeviceInterfaces([surf], device="cpu", z_platform=tr.positions[:, 2].min())
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        out, drop, bins, _ = dv.refracted_field(fac, trs, ifc, [1.0, 3.15], [0.0, 10.0],
                                                precision=precision, pair_bins=True,
                                                extra_flags=extra, **kw)
    o = out.numpy()
    ref_kw = dict(kw)
    if rough:
        ref_kw["roughness"] = (0.5, 20.0, ph, 24, "exponential")
    return (o[..., 0] + 1j * o[..., 1], drop.numpy(), fac.unpermute_pairs(bins.numpy()),
            (tr, surf, bed, ref_kw))


@pytest.mark.parametrize("rough", [False, True])
def test_refracted_fp64_kernel_vs_oracle(sim, rough):
    out, drop, bins, (tr, surf, bed, kw) = _refr(sim, "fp64", rough=rough)
    ref, dref, bref, vref = om.refracted_cluttergram(tr.positions, tr.u_ct, bed, [surf],
                                                     [1.0, 3.15], [0.0, 10.0],
                                                     return_bins=True, **kw)
    np.testing.assert_array_equal(bins, np.where(vref, bref, _lib.BIN_INVALID))
    assert np.linalg.norm(out - ref) <= 1e-9 * np.linalg.norm(ref)
    np.testing.assert_allclose(drop, dref, rtol=1e-9, atol=1e-30)


def test_mixed_redo_pass_reproduces_fp64(sim):
    """Every pair forced through the deferred float64 pass: flags, lane masks
    and the masked float64 trace loop give the float64 kernel's result."""
    o64, d64, b64, _ = _refr(sim, "fp64", rough=True)
    orr, drr, brr, _ = _refr(sim, "mixed", extra=_lib.SS_REDO_ALL, rough=True)
    om_, dm, bm, _ = _refr(sim, "mixed", rough=True)
    np.testing.assert_array_equal(brr, b64)
    np.testing.assert_array_equal(bm, b64)
    assert np.linalg.norm(orr - o64) <= 1e-12 * np.linalg.norm(o64)
    np.testing.assert_allclose(drr, d64, rtol=1e-12, atol=1e-30)
    pk = np.abs(o64).max(axis=-1, keepdims=True)
    m = np.abs(o64) > pk * 10 ** (-30 / 20)
    assert np.abs(20 * np.log10(np.abs(om_[m]) / np.abs(o64[m]))).max() <= 0.05
    assert np.abs(np.angle(om_[m] / o64[m])).max() * 1e3 <= 2.0


def test_surface_fp64_kernel_vs_oracle(sim):
    wl = W.config1(n=24, n_traces=5, n_samples=512, seed=11)
    f, tr = wl.layers[0].facets, wl.track
    kw = dict(k=wl.k0, gamma=wl.layers[0].gamma, t0=wl.t0, dt=wl.dt,
              n_samples=wl.n_samples, c=wl.c)
    ref, dref, bref, _ = oc.coherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals,
                                                 f.areas, f.e1, f.e2, return_bins=True, **kw)
    from radar_sounder_simulator_amd.kernels import _reference_range
    with _lib.use(sim):
        fac = dv.DeviceFacets(f.centers, f.normals, f.areas, f.e1, f.e2, device="cpu")
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, _reference_range(tr.positions, f.centers),
                              None, device="cpu")
        out, drop, bins, _ = dv.surface_field(fac, trs, precision="fp64", pair_bins=True, **kw)
    o = out.numpy()
    o = o[..., 0] + 1j * o[..., 1]
    np.testing.assert_array_equal(fac.unpermute_pairs(bins.numpy()), bref)
    assert np.linalg.norm(o - ref) <= 1e-9 * np.linalg.norm(ref)


def _flat_stack(seed=4, n=16):
    """Rough surface over three horizontal planes and a rough bed."""
    rng = np.random.default_rng(seed)
    dx = 25.0
    surf = grid_facets(3.0 * gaussian_surface(n + 1, dx, 100.0, rng), dx)
    planes = [grid_facets(np.full((n + 1, n + 1), z), dx) for z in (-30.0, -80.0, -150.0)]
    bed = grid_facets(-300.0 + 30.0 * gaussian_surface(n + 1, dx, 200.0, rng), dx)
    tr = straight_track(5, 40.0, 3000.0)
    return surf, planes, bed, tr


@pytest.mark.parametrize("pad_to", [None, 8])
@pytest.mark.parametrize("rough", [False, True])
def test_flat_stack_joint_kernel_vs_oracle(sim, pad_to, rough):
    """refraction code 2 (the joint solve specialised to horizontal planes
    below the surface) against the oracle's general joint path, with and
    without the no-op pads; the general joint kernel on the same stack."""
    surf, planes, bed, tr = _flat_stack()
    crossed = [surf] + planes
    eps, att = [1.0, 1.6, 2.1, 2.6, 3.15], [0.0, 5.0, 5.0, 5.0, 10.0]
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    ph = sig = None
    if rough:
        ph = speckle_phasors(len(bed), 5)
        sig = [0.2, 0.05, 0.05, 0.05]
        kw.update(roughness=(0.5, 20.0, 24, "exponential"), d_phi_area=True)
    outs = {}
    with _lib.use(sim):
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2,
                              phasors=ph, device="cpu")
        ifc = dv.DeviceInterfaces(crossed, sig, device="cpu", pad_to=pad_to or len(crossed),
                                  z_platform=tr.positions[:, 2].min())
        assert ifc.flat_below and ifc.horizontal == [False, True, True, True]
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        for flat in (True, False):
            out, drop, bins, _ = dv.refracted_field(fac, trs, ifc, eps, att, precision="fp64",
                                                    pair_bins=True, refraction="joint",
                                                    flat_stack=flat, **kw)
            o = out.numpy()
            outs[flat] = (o[..., 0] + 1j * o[..., 1], drop.numpy(),
                          fac.unpermute_pairs(bins.numpy()))
    ref_kw = dict(kw)
    if rough:
        ref_kw["roughness"] = (0.5, 20.0, ph, 24, "exponential")
    ref, dref, bref, vref = om.refracted_cluttergram(
        tr.positions, tr.u_ct, bed, crossed, eps, att, refraction="joint", pad_to=pad_to,
        crossed_sigma=sig, return_bins=True, **ref_kw)
    assert vref.mean() > 0.9
    for flat in (True, False):
        out, drop, bins = outs[flat]
        np.testing.assert_array_equal(bins, np.where(vref, bref, _lib.BIN_INVALID))
        assert np.linalg.norm(out - ref) <= 1e-9 * np.linalg.norm(ref)
        np.testing.assert_allclose(drop, dref, rtol=1e-9, atol=1e-30)
    # mixed tier (float64 flat solve, mixed radiometry) and its float64 redo
    with _lib.use(sim):
        for extra in (0, _lib.SS_REDO_ALL):
            out, drop, bins, _ = dv.refracted_field(fac, trs, ifc, eps, att, precision="mixed",
                                                    pair_bins=True, refraction="joint",
                                                    extra_flags=extra, flat_stack=True, **kw)
            o = out.numpy()
            o = o[..., 0] + 1j * o[..., 1]
            np.testing.assert_array_equal(fac.unpermute_pairs(bins.numpy()), outs[True][2])
            if extra:
                assert np.linalg.norm(o - outs[True][0]) <= 1e-12 * np.linalg.norm(ref)
            else:
                pk = np.abs(ref).max(axis=-1, keepdims=True)
                m = np.abs(ref) > pk * 10 ** (-30 / 20)
                assert np.abs(20 * np.log10(np.abs(o[m]) / np.abs(ref[m]))).max() <= 0.05
                assert np.abs(np.angle(o[m] / ref[m])).max() * 1e3 <= 2.0


def test_deep_flat_stack_kernel_vs_flat_oracle(sim):
    """41 interfaces (deeper than the general joint kernel takes): the
    flat-stack kernel against the oracle of the flat-stack form.  The rough
    surface dips below the first plane (-5 m) in places: those lanes are
    invalid and their power is booked in dropped from the unsolved
    polyline."""
    surf, _, bed, tr = _flat_stack(seed=5, n=20)
    n = surf.grid_shape[0]
    crossed = [surf] + [grid_facets(np.full((n + 1, n + 1), z), 25.0)
                        for z in -5.0 * np.arange(1, 41)]
    eps = list(np.concatenate([[1.0], np.linspace(1.3, 3.1, 40), [3.15]]))
    att = [0.0] + [5.0] * 40 + [10.0]
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    with _lib.use(sim):
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2, device="cpu")
        ifc = dv.DeviceInterfaces(crossed, device="cpu", pad_to=len(crossed),
                                  z_platform=tr.positions[:, 2].min())
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        out, drop, bins, _ = dv.refracted_field(fac, trs, ifc, eps, att, precision="fp64",
                                                pair_bins=True, refraction="joint",
                                                flat_stack=True, **kw)
    o = out.numpy()
    o = o[..., 0] + 1j * o[..., 1]
    ref, dref, bref, vref = om.refracted_cluttergram(tr.positions, tr.u_ct, bed, crossed, eps,
                                                     att, return_bins=True, refraction="joint",
                                                     flat_stack=True, **kw)
    assert 0.9 < vref.mean() < 1.0
    np.testing.assert_array_equal(fac.unpermute_pairs(bins.numpy()),
                                  np.where(vref, bref, _lib.BIN_INVALID))
    assert np.linalg.norm(o - ref) <= 1e-9 * np.linalg.norm(ref)
    np.testing.assert_allclose(drop.numpy(), dref, rtol=1e-9, atol=1e-30)
    # mixed tier: lanes with no feasible start step settle without the
    # float64 redo and book their dropped power from the start path
    with _lib.use(sim):
        om_, dm, bm, _ = dv.refracted_field(fac, trs, ifc, eps, att, precision="mixed",
                                            pair_bins=True, refraction="joint",
                                            flat_stack=True, **kw)
    np.testing.assert_array_equal(fac.unpermute_pairs(bm.numpy()),
                                  np.where(vref, bref, _lib.BIN_INVALID))
    np.testing.assert_allclose(dm.numpy(), dref, rtol=1e-4, atol=1e-30)


def test_multilayer_scene_layers_vs_oracle(sim):
    """The multilayer workload (rough surface, 20 firn sublayers, three
    horizons, rough bed; every interface a target) through the resident
    pipeline: each refracted layer against the oracle (sequential for the
    first sublayer, the flat-stack joint form below it, padded as the
    reference's simulate() pads), float64."""
    wl = W.config3(nx=16, ny=8, n_traces=2)
    with _lib.use(sim):
        rs = ResidentScene(wl, device="cpu")
        for j in range(len(rs.layers)):
            rs.run_layer(j, "fp64")
        raw = rs.raw.numpy().copy()
    tr = wl.track
    assert len(wl.layers) == 25 and wl.layers[-1].pad_to == 24
    for j, L in enumerate(wl.layers[1:], start=1):
        f = L.facets
        ref, _ = om.refracted_cluttergram(
            tr.positions, tr.u_ct, f, L.crossed, L.eps_leg, L.att_leg, mode="coherent",
            t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c, gamma=L.gamma, k0=wl.k0,
            pattern=wl.pattern, refraction=L.refraction, pad_to=L.pad_to,
            flat_stack=L.flat_stack)
        o = raw[j][..., 0] + 1j * raw[j][..., 1]
        assert np.linalg.norm(ref) > 0, L.name
        assert np.linalg.norm(o - ref) <= 1e-9 * np.linalg.norm(ref), L.name


def test_sequential_chain_over_horizontal_planes_vs_oracle(sim):
    """Horizontal planes are stored as one cell; the sequential chain
    through them matches the oracle's (full-table) lookups bit for bit in
    its bins."""
    surf, planes, bed, tr = _flat_stack()
    crossed = [surf] + planes
    eps, att = [1.0, 1.6, 2.1, 2.6, 3.15], [0.0, 5.0, 5.0, 5.0, 10.0]
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    with _lib.use(sim):
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2, device="cpu")
        ifc = dv.DeviceInterfaces(crossed, device="cpu")
        assert ifc.tables.shape[0] == len(surf) + 3
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        out, drop, bins, _ = dv.refracted_field(fac, trs, ifc, eps, att, precision="fp64",
                                                pair_bins=True, **kw)
    o = out.numpy()
    o = o[..., 0] + 1j * o[..., 1]
    ref, dref, bref, vref = om.refracted_cluttergram(tr.positions, tr.u_ct, bed, crossed, eps,
                                                     att, return_bins=True, **kw)
    np.testing.assert_array_equal(fac.unpermute_pairs(bins.numpy()),
                                  np.where(vref, bref, _lib.BIN_INVALID))
    assert np.linalg.norm(o - ref) <= 1e-12 * np.linalg.norm(ref)


def _scenes():
    return [W.config2(n=9, n_traces=9, n_samples=4096, seed=3),
            W.config1(n=17, n_traces=5, n_samples=512, seed=8)]


def _run_single(sim, wls, precision):
    outs = []
    with _lib.use(sim):
        for w in wls:
            rs = ResidentScene(w, device="cpu")
            for j in range(len(rs.layers)):
                rs.run_layer(j, precision)
            outs.append(rs.raw.numpy().copy())
    return outs


def _run_batch(sim, wls, precision, tt, redo_all=False):
    with _lib.use(sim):
        scenes = [ResidentScene(w, device="cpu") for w in wls]
        jobs, keep = [], []
        for rs in scenes:
            j, k = rs.jobs(precision)
            jobs += j
            keep += k
        if redo_all:
            for jb in jobs:
                if jb.kind == 1:
                    jb.refracted.flags |= _lib.SS_REDO_ALL
        B.BatchRunner(device="cpu").launch(jobs, precision, trace_tile=tt, n_blocks=3)
        return [rs.raw.numpy().copy() for rs in scenes]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("tt", [4, 16])
def test_batched_launch_matches_single_launches(sim, tt):
    wls = _scenes()
    ref64 = _run_single(sim, wls, "fp64")
    for o, r in zip(_run_batch(sim, wls, "fp64", tt), ref64):
        assert all(np.linalg.norm(r[j]) > 0 for j in range(r.shape[0]))
        assert _rel(o, r) <= 1e-12
    for o, r in zip(_run_batch(sim, wls, "mixed", tt, redo_all=True), ref64):
        assert _rel(o[1:], r[1:]) <= 1e-12          # refracted layers: float64 redo
        assert _rel(o[:1], r[:1]) <= 1e-4           # surface layer: mixed tier


@pytest.mark.parametrize("n,m", [(300, 20), (1000, 130)])
def test_fft_compression_kernel_vs_oracle(sim, n, m):
    """Overlap-save FFT compression (several blocks per row) against the
    oracle's FFT convolution, and against the direct-summation kernel."""
    from oracle import waveform as ow
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    p = rng.standard_normal(2 * m + 1) + 1j * rng.standard_normal(2 * m + 1)
    ref = ow.convolve_fast_time(x, p, m)
    t = torch.as_tensor(np.stack([x.real, x.imag], -1))
    with _lib.use(sim):
        y, ps = dv.range_compress(t, p, m, method="fft")
        d, _ = dv.range_compress(t, p, m, method="direct")
    L = ps.nfft - 2 * m
    assert ps.nfft >= 256 and (n + L - 1) // L >= 2        # several blocks per row
    y = y.numpy()
    y = y[..., 0] + 1j * y[..., 1]
    d = d.numpy()
    d = d[..., 0] + 1j * d[..., 1]
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.linalg.norm(d - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_device_resident_step_matches_batched_launch(sim, precision):
    """ss_batch_step reads the job count and table from device memory and
    sizes nothing on the host: the same results as ss_batch_field, for a
    repacked table of a different size too."""
    wls = _scenes()
    ref = _run_batch(sim, wls, precision, 8)
    tol = 1e-12 if precision == "fp64" else 1e-6
    with _lib.use(sim):
        scenes = [ResidentScene(w, device="cpu") for w in wls]
        jobs, keep = [], []
        for rs in scenes:
            j, k = rs.jobs(precision)
            jobs += j
            keep += k
        sg = B.StepGraph("cpu", precision, 8, feature_class=1, max_jobs=8, max_items=4096,
                         n_blocks=3, capture=False)
        assert sg.fits(jobs)[0]
        for subset in (jobs, jobs[:2]):          # second launch: fewer jobs
            for rs in scenes:
                rs.raw.fill_(7.0)
            sg.launch(subset)
            n_layers = [len(rs.layers) for rs in scenes]
            done = 0
            for rs, r, nl in zip(scenes, ref, n_layers):
                for j in range(nl):
                    if done < len(subset):
                        assert _rel(rs.raw[j].numpy(), r[j]) <= tol
                    else:
                        assert np.all(rs.raw[j].numpy() == 7.0)   # not in this table
                    done += 1


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("n_samples,r0", [(600, 700.0), (80, 820.0)])
def test_incoherent_kernel_vs_oracle(sim, split, n_samples, r0):
    """Incoherent kernel (both tiers) against oracle/incoherent.py: bins
    exact; power and dropped to 1e-12 (float64) / 1e-6 (mixed); the short
    window cuts most pairs into dropped."""
    from oracle import incoherent as oi
    rng = np.random.default_rng(3)
    f = grid_facets(20.0 * gaussian_surface(41, 25.0, 150.0, rng), 25.0)
    tr = straight_track(7, 60.0, 800.0)
    kw = dict(t0=2 * r0 / C, dt=1 / 48e6, n_samples=n_samples, c=C, split_sides=split)
    ref, dref, bref = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals,
                                                f.areas, return_bins=True, **kw)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        with _lib.use(sim):
            fac = dv.DeviceFacets(f.centers, f.normals, f.areas, f.e1, f.e2, device="cpu")
            trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
            out, drop, bins, _ = dv.incoherent_power(fac, trs, precision=prec, pair_bins=True,
                                                     **kw)
        o = out.numpy()
        if split:
            o = o.reshape(-1, 2, n_samples).transpose(0, 2, 1)
        np.testing.assert_array_equal(fac.unpermute_pairs(bins.numpy()), bref)
        assert np.linalg.norm(o - ref) <= tol * np.linalg.norm(ref)
        np.testing.assert_allclose(drop.numpy(), dref, rtol=tol, atol=1e-300)


def test_incoherent_refracted_dropped_is_power(sim):
    """Regression: an out-of-window incoherent refracted pair books its power
    (not its square) to dropped."""
    surf, bed, tr = _bed()
    kw = dict(mode="incoherent", t0=2 * 3200 / C, dt=1 / 48e6, n_samples=40, c=C,
              gamma=0.0, k0=None)
    with _lib.use(sim):
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2, device="cpu")
        ifc = dv.DeviceInterfaces([surf], device="cpu", z_platform=tr.positions[:, 2].min())
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        out, drop, _ = dv.refracted_field(fac, trs, ifc, [1.0, 3.15], [0.0, 10.0],
                                          precision="fp64", **kw)
    ref, dref = om.refracted_cluttergram(tr.positions, tr.u_ct, bed, [surf], [1.0, 3.15],
                                         [0.0, 10.0], **kw)
    assert (dref > 0).all()
    np.testing.assert_allclose(drop.numpy(), dref, rtol=1e-9)
    np.testing.assert_allclose(out.numpy()[..., 0], ref, rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("n,stride", [(4, 4), (4, 1), (1, 1), (9, 9)])
def test_slow_time_boxcar_vs_oracle(sim, n, stride):
    """presum / unfocused / multilook windows of ss_slow_time_boxcar against
    oracle/processing.py's weight matrices (complex as 2 real columns)."""
    from oracle import processing as opr
    rng = np.random.default_rng(8)
    T = 37
    x = rng.standard_normal((T, 30)) + 1j * rng.standard_normal((T, 30))
    W = opr.boxcar_weights(T, n, stride)
    ref = opr.apply_weights(W, x)
    with _lib.use(sim):
        xt = torch.as_tensor(np.stack([x.real, x.imag], -1))
        y = dv.slow_time_boxcar(xt, n, stride, W.shape[0]).numpy()
        m = dv.slow_time_boxcar(torch.as_tensor(np.abs(x) ** 2), n, stride, W.shape[0],
                                scale=1.0 / n).numpy()
    np.testing.assert_allclose(y[..., 0] + 1j * y[..., 1], ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m, opr.apply_weights(W, np.abs(x) ** 2, average=True),
                               rtol=1e-13)


@pytest.mark.parametrize("window", ["hann", "none"])
def test_focused_sar_kernel_vs_oracle(sim, window):
    """Backprojection kernel, both tiers, against oracle/processing.py on a
    point-target history plus noise: float64 to 1e-12, mixed to 1e-5."""
    from oracle import processing as opr
    rng = np.random.default_rng(9)
    T, nb, dx, H = 41, 96, 2.0, 600.0
    lam = C / 150e6
    k = 2 * np.pi / lam
    B = 15e6
    dt = 1 / (4 * B)
    t0 = 2 * H / C - 30 * dt
    xs = (np.arange(T) - T // 2) * dx
    R = np.sqrt(H * H + xs * xs)
    twtt = t0 + dt * np.arange(nb)
    F = np.sinc(B * (twtt[None, :] - 2 * R[:, None] / C)) * np.exp(-2j * k * R)[:, None]
    F = F + 0.05 * (rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape))
    s = opr.along_track(np.column_stack([xs, np.zeros(T), np.full(T, H)]))
    ref = opr.focused_sar(F, s, twtt, c=C, k=k, aperture_m=30.0, window=window)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        with _lib.use(sim):
            out = dv.focused_sar(torch.as_tensor(np.stack([F.real, F.imag], -1)),
                                 torch.as_tensor(twtt), torch.as_tensor(s), c=C, k=k,
                                 aperture_m=30.0, window=window, precision=prec).numpy()
        o = out[..., 0] + 1j * out[..., 1]
        assert np.linalg.norm(o - ref) <= tol * np.linalg.norm(ref), prec


def test_noalloc_entry_matches_workspace_entry(sim):
    """ss_refracted_field_noalloc (the library's scratch array) against
    ss_refracted_field with a caller workspace, every pair through the redo
    record, two calls in a row, the second smaller."""
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    for n, seed in ((12, 5), (7, 6)):
        surf, bed, tr = _bed(seed=seed, n=n)
        res = {}
        with _lib.use(sim):
            fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2,
                                  device="cpu")
            ifc = dv.DeviceInterfaces([surf], device="cpu")
            trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
            for noalloc in (False, True):
                out, drop, bins, _ = dv.refracted_field(
                    fac, trs, ifc, [1.0, 3.15], [0.0, 10.0], precision="mixed",
                    pair_bins=True, extra_flags=_lib.SS_REDO_ALL, noalloc=noalloc, **kw)
                res[noalloc] = (out.numpy().copy(), drop.numpy().copy(), bins.numpy().copy())
        np.testing.assert_array_equal(res[True][2], res[False][2])
        assert np.linalg.norm(res[False][0]) > 0
        np.testing.assert_allclose(res[True][0], res[False][0], rtol=0,
                                   atol=1e-12 * np.abs(res[False][0]).max())
        np.testing.assert_allclose(res[True][1], res[False][1], rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("flat", [False, True])
def test_batched_joint_jobs_match_single_launches(sim, flat):
    """Batched jobs of the joint solver classes (general joint solve, and its
    flat-stack form) against one launch per layer: float64 to rounding, the
    mixed tier with every refracted pair through the float64 redo to
    rounding of the float64 single launches, and the plain mixed tier to its
    summation order."""
    wls = [W.config3(nx=16, ny=8, n_traces=3, n_firn=3, horizons=(500.0,), flat_stack=flat),
           W.config2(n=9, n_traces=5, n_samples=4096, seed=4)]
    assert any(L.refraction == "joint" for L in wls[0].layers)
    ref64 = _run_single(sim, wls, "fp64")
    for o, r in zip(_run_batch(sim, wls, "fp64", 8), ref64):
        assert all(np.linalg.norm(r[j]) > 0 for j in range(r.shape[0]))
        assert _rel(o, r) <= 1e-12
    for o, r in zip(_run_batch(sim, wls, "mixed", 8, redo_all=True), ref64):
        assert _rel(o[1:], r[1:]) <= 1e-12
    refm = _run_single(sim, wls, "mixed")
    for o, r in zip(_run_batch(sim, wls, "mixed", 4), refm):
        assert _rel(o, r) <= 1e-6


def test_batch_step_overflow_word(sim):
    """A mixed step whose refracted items exceed the redo record sets the
    workspace's overflow word (ss_batch_overflowed); a step within it
    clears it."""
    import ctypes
    wls = _scenes()
    with _lib.use(sim):
        scenes = [ResidentScene(w, device="cpu") for w in wls]
        jobs, keep = [], []
        for rs in scenes:
            j, k = rs.jobs("mixed")
            jobs += j
            keep += k
        _, refr, _ = B.pack_info(jobs, 8)
        lib = _lib.load()
        flag = ctypes.c_int(-1)
        for cap, want in ((max(1, refr // 2), 1), (refr, 0)):
            sg = B.StepGraph("cpu", "mixed", 8, feature_class=1, max_jobs=8, max_items=cap,
                             n_blocks=3, capture=False)
            # bypass launch()'s own capacity check: pack and step directly
            arr = (_lib.Job * len(jobs))(*jobs)
            rc = lib.ss_batch_pack(arr, len(jobs), 8, ctypes.c_void_p(sg.table_host.data_ptr()),
                                   sg.table_host.numel(), None, None)
            _lib.check(rc, "ss_batch_pack")
            sg._enqueue()
            rc = lib.ss_batch_overflowed(ctypes.c_void_p(sg.workspace.data_ptr()),
                                         ctypes.byref(flag), None)
            _lib.check(rc, "ss_batch_overflowed")
            assert flag.value == want, (cap, refr)

"""GPU: along-track processing (presum, unfocused SAR, multilook, focused
SAR) against oracle/processing.py, and the reference's own processing
checks (tests/test_processing.py:56-311 of soundersim) through the
package API on SimResult."""

import json
import warnings

import numpy as np
import pytest
import torch

from oracle import processing as op
from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd import processing as proc
from radar_sounder_simulator_amd.output import SimResult

pytestmark = pytest.mark.gpu
C = 299792458.0
F0 = 195e6
LAM = C / F0


def _result(field, twtt, x, z=1000.0, mode="coherent"):
    T = field.shape[0]
    pos = np.column_stack([np.asarray(x, float), np.zeros(T), np.full(T, z)])
    cfg = json.dumps({"radar": {"c": C, "f0": F0}})
    res = SimResult(field=np.asarray(field, np.complex64) if mode == "coherent" else None,
                    dropped_power=np.zeros(T, np.float32),
                    nadir_twtt=np.full(T, float(twtt[0])), twtt=np.asarray(twtt, float),
                    positions=pos, attrs={"mode": mode, "config": cfg})
    if mode != "coherent":
        res.power_data = (np.abs(field) ** 2).astype(np.float32)
    return res


def _point_target(H=200.0, aperture=80.0, spacing=0.25, extra=10.0, dt=2e-9, bandwidth=30e6):
    """A point scatterer below a straight level track: hann-weighted
    compressed-pulse envelope at the exact delay, carrier phase
    exp(-2 j k r) (the reference's construction, tests/test_processing.py:
    159-186)."""
    k = 2.0 * np.pi / LAM
    s = np.arange(-(aperture / 2 + extra), aperture / 2 + extra + spacing / 2, spacing)
    r = np.sqrt(s ** 2 + H ** 2)
    t0 = 2.0 * (H - 30.0) / C
    nb = int(np.ceil((2.0 * r.max() / C + 60.0 * 2.0 / C - t0) / dt)) + 4
    twtt = t0 + np.arange(nb) * dt
    x = bandwidth * (twtt[None, :] - 2.0 * r[:, None] / C)
    p = (0.5 * np.sinc(x) + 0.25 * (np.sinc(x - 1) + np.sinc(x + 1))) / 0.5
    return p * np.exp(-2j * k * r[:, None]), s, twtt, H


@pytest.mark.parametrize("window", ["hann", "none"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_focus_kernel_vs_oracle(window, precision):
    print(f"case processing/focus/{window}/{precision}", flush=True)
    F, s, twtt, H = _point_target()
    rng = np.random.default_rng(4)
    F = F + 0.02 * (rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape))
    k = 2 * np.pi / LAM
    ref = op.focused_sar(F, s - s[0], twtt, c=C, k=k, aperture_m=80.0, window=window)
    d = torch.device("cuda")
    out = dv.focused_sar(torch.as_tensor(np.stack([F.real, F.imag], -1), device=d),
                         torch.as_tensor(twtt, device=d), torch.as_tensor(s - s[0], device=d),
                         c=C, k=k, aperture_m=80.0, window=window, precision=precision)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    o = o[..., 0] + 1j * o[..., 1]
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert np.linalg.norm(o - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("n,stride", [(5, 5), (5, 1), (16, 16)])
def test_boxcar_kernel_vs_oracle(n, stride):
    print(f"case processing/boxcar/{n}/{stride}", flush=True)
    rng = np.random.default_rng(6)
    T = 203
    x = rng.standard_normal((T, 700, 2))
    W = op.boxcar_weights(T, n, stride)
    d = torch.device("cuda")
    y = dv.slow_time_boxcar(torch.as_tensor(x, device=d), n, stride, W.shape[0])
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), op.apply_weights(W, x), rtol=1e-12, atol=1e-12)


def test_presum_unfocused_point_target_gain_and_conventions():
    """presum gains n on an in-phase target, equals every n-th unfocused
    output, averages positions, records the step, refuses incoherent."""
    print("case processing/presum", flush=True)
    T, n = 40, 4
    field = np.ones((T, 3), np.complex64) * np.exp(1j * 0.3)
    res = _result(field, twtt=[1e-6, 1.01e-6, 1.02e-6], x=np.arange(T) * 2.0)
    p = proc.presum(res, n)
    u = proc.unfocused_sar(res, n_traces=n)
    assert p.field.shape == (T // n, 3) and u.field.shape == (T - n + 1, 3)
    np.testing.assert_allclose(np.abs(p.field), n, rtol=1e-6)
    np.testing.assert_array_equal(p.field, u.field[::n])
    np.testing.assert_allclose(p.positions[:, 0], np.arange(T // n) * 2.0 * n + 3.0)
    step = json.loads(p.attrs["processing"])[-1]
    assert step["op"] == "presum" and step["n"] == n and step["n_traces_out"] == T // n
    u8 = proc.unfocused_sar(res, aperture_m=8.0)   # spacing 2 m -> 4 traces
    assert json.loads(u8.attrs["processing"])[-1]["n_traces"] == 4
    inc = _result(field, twtt=[1e-6, 1.01e-6, 1.02e-6], x=np.arange(T) * 2.0,
                  mode="incoherent")
    with pytest.raises(ValueError, match="multilook"):
        proc.presum(inc, 2)
    m = proc.multilook(inc, 8)
    assert m.field is None and m.power.shape == (T // 8, 3)
    np.testing.assert_allclose(m.power, 1.0, rtol=1e-6)


def test_multilook_speckle_contrast():
    print("case processing/multilook", flush=True)
    rng = np.random.default_rng(3)
    T, n = 4096, 16
    z = (rng.standard_normal((T, 256)) + 1j * rng.standard_normal((T, 256))) / np.sqrt(2)
    res = _result(z, twtt=1e-6 + 1e-8 * np.arange(256), x=np.arange(T) * 1.0)
    c1 = res.power.std() / res.power.mean()
    m = proc.multilook(res, n)
    cn = m.power.std() / m.power.mean()
    assert abs(c1 - 1) < 0.02 and abs(cn * np.sqrt(n) - 1) < 0.03


def _width_3db(amp, coord):
    p = 20 * np.log10(np.maximum(amp, 1e-12) / amp.max())
    i0 = int(np.argmax(p))

    def cross(rng_):
        prev = i0
        for i in rng_:
            if p[i] <= -3.0:
                f = (-3.0 - p[prev]) / (p[i] - p[prev])
                return coord[prev] + f * (coord[i] - coord[prev])
            prev = i
        return None
    return cross(range(i0 + 1, len(p))) - cross(range(i0 - 1, -1, -1))


def test_focused_sar_resolution_gain_and_guards():
    """tests/test_processing.py:209-279 of the reference through the package
    API: rectangular-aperture width within 15 % of lambda r / (2 L), gain >
    0.7 of the aperture trace count, sidelobes < -10 dB, focused beats
    unfocused, 2-D field required, the Doppler guard warns when sparse and
    stays quiet when dense."""
    print("case processing/focused", flush=True)
    L = 80.0
    F, s, twtt, H = _point_target(aperture=L)
    res = _result(F, twtt, s, z=H)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        f = proc.focused_sar(res, aperture_m=L, window="none")
    assert not [w for w in caught if "alias" in str(w.message)]
    b0 = int(round((2.0 * H / C - twtt[0]) / (twtt[1] - twtt[0])))
    amp = np.abs(f.field[:, b0])
    pred = LAM * H / (2.0 * L)
    width = _width_3db(amp, s)
    assert 0.85 <= width / pred <= 1.15, (width, pred)
    n_ap = int(np.sum(np.abs(s - s[np.argmax(amp)]) <= L / 2))
    assert amp.max() / n_ap > 0.7
    prof = 20 * np.log10(amp / amp.max())
    assert prof[np.abs(s - s[np.argmax(amp)]) > 2.0 * pred].max() < -10.0
    u = proc.unfocused_sar(res, n_traces=n_ap)
    assert np.abs(u.field).max() < 0.5 * amp.max()
    assert json.loads(f.attrs["processing"])[-1]["op"] == "focused_sar"
    three = _result(F, twtt, s, z=H)
    three.field = np.stack([three.field, three.field], -1)
    with pytest.raises(ValueError, match="2-D"):
        proc.focused_sar(three, aperture_m=L)
    sparse = _result(F[::40], twtt, s[::40], z=H)
    with pytest.warns(UserWarning, match="alias"):
        proc.focused_sar(sparse, aperture_m=L)

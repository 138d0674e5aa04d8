"""Simulation output container (the reference's ``output.py:47-197`` Dataset
without xarray): ``field`` (T, n[, 2][, L]) complex, ``power`` = |field|^2,
``dropped_power``, per-layer ``nadir_twtt``, the fast-time axis and
provenance attributes; ``combine`` sums layers (coherent fields add before
squaring); ``save``/``load`` round-trip through a NumPy ``.npz`` with a JSON
attribute record (NetCDF needs h5netcdf, which this library does not use).

The reference's ``first_return_*`` diagnostics (``output.py:77-82``, a
(T, F, 3) float64 host array there, ~50 GB at the 2.1 M-facet scale) are
computed on the device (``kernels.trace_diagnostics``):
``first_return_twtt`` and the facet centre ``first_return_xyz`` in the
local frame (the reference converts it to lat/lon; geodesy is not part of
this library).
"""

import json
from dataclasses import dataclass, field

import numpy as np

VERSION = "0.1"


@dataclass
class SimResult:
    field: np.ndarray
    dropped_power: np.ndarray
    nadir_twtt: np.ndarray
    twtt: np.ndarray
    positions: np.ndarray
    layer: list = None
    attrs: dict = field(default_factory=dict)
    first_return_twtt: np.ndarray = None
    first_return_xyz: np.ndarray = None

    @property
    def power(self):
        return np.abs(self.field) ** 2

    @classmethod
    def build(cls, *, field, dropped, nadir_twtt, t0, dt, positions, layers=None,
              config=None, first_return_twtt=None, first_return_xyz=None):
        n = field.shape[1]
        attrs = {"soundersim_amd_version": VERSION, "mode": "coherent"}
        if config is not None:
            attrs["config"] = config.model_dump_json()
        nad = nadir_twtt if layers is not None else nadir_twtt[:, 0]
        return cls(field, dropped, nad, t0 + dt * np.arange(n), np.asarray(positions),
                   list(layers) if layers is not None else None, attrs,
                   first_return_twtt, first_return_xyz)

    def sel_layer(self, name):
        j = self.layer.index(name)
        return self.field[..., j]


def combine(res, dim="layer"):
    """Total power over layers: |sum of fields|^2 (coherent)."""
    if dim != "layer" or res.layer is None:
        raise ValueError("combine over 'layer' needs a multilayer result")
    return np.abs(res.field.sum(axis=-1)) ** 2


def save(res, path):
    extra = {}
    if res.first_return_twtt is not None:
        extra = dict(first_return_twtt=res.first_return_twtt,
                     first_return_xyz=res.first_return_xyz)
    np.savez(path, field=res.field, dropped_power=res.dropped_power,
             nadir_twtt=res.nadir_twtt, twtt=res.twtt, positions=res.positions,
             meta=np.frombuffer(json.dumps({"layer": res.layer, "attrs": res.attrs}).encode(),
                                dtype=np.uint8), **extra)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(z["meta"].tobytes().decode())
        fr = ((z["first_return_twtt"], z["first_return_xyz"]) if "first_return_twtt" in z
              else (None, None))
        return SimResult(z["field"], z["dropped_power"], z["nadir_twtt"], z["twtt"],
                         z["positions"], meta["layer"], meta["attrs"], *fr)

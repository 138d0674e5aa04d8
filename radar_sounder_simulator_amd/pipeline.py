"""Resident scene: one workload's facets, interfaces and traces kept in HBM,
evaluated end to end (every target layer + range compression) on a stream.

This is the simulate() hot path of the reference (``simulate.py:263-395``:
surface through the coherent kernel, each deeper interface through the
refracted kernel, stacked per layer, then ``apply_waveform``) without the
host-side scene construction, with all buffers allocated once.  A workload
whose every layer is a lattice is built in HBM from its heights (or, for
seeded lattices, from the seeds alone: ``device.build_lattices``).
"""

import numpy as np
import torch

from . import device as dv
from .kernels import _reference_range
from .waveform import compressed_pulse


class ResidentScene:
    """Workload uploaded to ``device``; ``run(precision)`` enqueues one pass."""

    def __init__(self, wl, device=None):
        self.wl = wl
        self.device = dv._dev(device)
        tr = wl.track
        pv = None if wl.pattern is None else wl.pattern[1]
        self.layers = []
        self.built = None
        refr = [L for L in wl.layers if L.crossed]
        lat_of = {id(L.facets): L.lattice for L in wl.layers if L.lattice is not None}
        if (all(L.lattice is not None for L in wl.layers)
                and all(id(c) in lat_of for L in refr for c in L.crossed)):
            facs, stacks, center_mean = self._build_on_device(wl, lat_of, refr)
            d = tr.positions - center_mean
            r_ref = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        else:
            # host-built facets; every layer's crossed stack over one shared table
            facs = []
            for L in wl.layers:
                ph = None if L.roughness is None else L.roughness[2]
                g = np.asarray(L.gamma, np.float64)
                facs.append(dv.DeviceFacets.from_facets(L.facets, gamma=g if g.ndim else None,
                                                        phasors=ph, device=self.device))
            stacks = iter(dv.DeviceInterfaces.stacks(
                [L.crossed for L in refr], None, [L.pad_to for L in refr],
                tr.positions[:, 2].min(), self.device) if refr else [])
            r_ref = _reference_range(tr.positions, wl.layers[0].facets.centers)
        self.traces = dv.DeviceTraces(tr.positions, tr.u_ct, r_ref, pv, device=self.device)
        for L, fac in zip(wl.layers, facs):
            ifc = next(stacks) if L.crossed else None
            rough = (None if L.roughness is None else
                     (L.roughness[0], L.roughness[1], L.roughness[3],
                      L.roughness[4] if len(L.roughness) > 4 else "gaussian"))
            self.layers.append((L, fac, ifc, rough))
        T = tr.positions.shape[0]
        n_seg = (2 if wl.split_sides else 1) * wl.n_samples
        nl = len(wl.layers)
        self.raw = torch.zeros((nl, T, n_seg, 2), dtype=torch.float64, device=self.device)
        self.out = torch.zeros_like(self.raw)
        self.dropped = torch.zeros((nl, T), dtype=torch.float64, device=self.device)
        self.pulse = None
        if wl.waveform is not None and wl.waveform.kind != "delta":
            p, m = compressed_pulse(wl.waveform.bandwidth, wl.waveform.pulse_length,
                                    wl.dt, wl.waveform.window, wl.waveform.construction,
                                    wl.waveform.tukey_alpha)
            self.pulse = (p, m)
            self.pulse_prep = dv.PulseSpectrum(p, m, wl.n_samples, self.device)
        self._keep = []
        # float64-redo workspaces of the mixed refracted launches, one per layer
        self.workspace = [
            None if ifc is None else torch.empty(
                max(1, dv.refracted_workspace_bytes("mixed", fac.n, T)),
                dtype=torch.uint8, device=self.device)
            for _, fac, ifc, _ in self.layers]

    def _build_on_device(self, wl, lat_of, refr):
        """Every layer a lattice: heights drawn (seeded lattices) or uploaded,
        tessellated, tabulated and keyed phasors drawn in HBM
        (``device.build_lattices``); the crossed stacks, pads included, over
        the one shared table.  Returns (facets per layer, stack iterator,
        surface centre mean)."""
        tr = wl.track
        lats = [L.lattice for L in wl.layers]
        crossed = {id(lat_of[id(c)]) for L in refr for c in L.crossed}
        gammas, phasors = {}, {}
        for L in wl.layers:
            g = np.asarray(L.gamma, np.float64)
            if g.ndim:
                gammas[id(L.lattice)] = g
            if L.roughness is not None:
                phasors[id(L.lattice)] = L.roughness[2]
        n_pad = sum(max(0, (L.pad_to or 0) - len(L.crossed)) for L in refr)
        built, table = dv.build_lattices(lats, crossed, device=self.device, gammas=gammas,
                                         phasors=phasors, n_pad_rows=n_pad)
        stacks = iter(dv.DeviceInterfaces.lattice_stacks(
            [[lat_of[id(c)] for c in L.crossed] for L in refr], built, table,
            pad_row0=table.shape[0] - n_pad, pad_tos=[L.pad_to for L in refr],
            z_platform=tr.positions[:, 2].min(), device=self.device) if refr else [])
        self.built = built
        return ([built[id(L.lattice)].fac for L in wl.layers], stacks,
                built[id(lats[0])].center_mean)

    def pairs_per_layer(self):
        T = self.traces.n
        return [T * fac.n for _, fac, _, _ in self.layers]

    def nbytes(self):
        b = self.traces.buf.numel() * 8 + self.raw.numel() * 16 + self.dropped.numel() * 8
        b += sum(w.numel() for w in self.workspace if w is not None)
        tables = set()
        for _, fac, ifc, _ in self.layers:
            b += fac.nbytes()
            if ifc is not None:
                b += ifc.desc.numel() * 8
                if ifc.tables.data_ptr() not in tables:
                    tables.add(ifc.tables.data_ptr())
                    b += ifc.tables.numel() * 8
        return b

    def run_layer(self, j, precision, stream=None):
        wl = self.wl
        L, fac, ifc, rough = self.layers[j]
        if ifc is None:
            res = dv.surface_field(
                fac, self.traces, k=wl.k0, gamma=0.0 if fac.has_gamma else L.gamma,
                t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c,
                precision=precision, split_sides=wl.split_sides, pattern=wl.pattern,
                roughness=rough, taper_s=wl.taper_s, d_phi_area=wl.d_phi_area,
                out=self.raw[j], dropped=self.dropped[j], stream=stream)
        else:
            res = dv.refracted_field(
                fac, self.traces, ifc, L.eps_leg, L.att_leg, mode="coherent",
                t0=wl.t0, dt=wl.dt, n_samples=wl.n_samples, c=wl.c,
                gamma=0.0 if fac.has_gamma else L.gamma, k0=wl.k0,
                precision=precision, split_sides=wl.split_sides, pattern=wl.pattern,
                roughness=rough, taper_s=wl.taper_s, d_phi_area=wl.d_phi_area,
                out=self.raw[j], dropped=self.dropped[j], workspace=self.workspace[j],
                refraction=L.refraction, flat_stack=L.flat_stack, stream=stream)
        self._keep.append(res[-1])

    def compress(self, stream=None):
        if self.pulse is None:
            self.out.copy_(self.raw)
            return
        n = self.wl.n_samples
        dv.range_compress(self.raw.view(-1, n, 2), self.pulse_prep,
                          out=self.out.view(-1, n, 2), stream=stream)

    def run(self, precision, stream=None):
        self._keep = []
        for j in range(len(self.layers)):
            self.run_layer(j, precision, stream)
        self.compress(stream)
        return self.out

    def jobs(self, precision, out=None, dropped=None):
        """``ss_job`` records (one per layer, each refracted layer with its own
        solver: sequential chain, joint solve or its flat-stack form) writing
        into this scene's ``raw``/``dropped`` buffers (or the given ones), and
        the tensors that must outlive the launch."""
        from . import _lib
        wl = self.wl
        out = self.raw if out is None else out
        dropped = self.dropped if dropped is None else dropped
        jobs, keep = [], []
        cst = dv.constants(self.device)
        for j, (L, fac, ifc, rough) in enumerate(self.layers):
            jb = _lib.Job()
            jb.facet_stride, jb.n_facets = fac.stride, fac.n
            jb.trace_stride, jb.n_traces = self.traces.stride, self.traces.n
            jb.facets, jb.traces = fac.buf.data_ptr(), self.traces.buf.data_ptr()
            jb.field, jb.dropped = out[j].data_ptr(), dropped[j].data_ptr()
            jb.constants = cst.data_ptr()
            g = 0.0 if fac.has_gamma else L.gamma
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
                keep.append(legs)
            jb.pattern_table = 0 if table is None else table.data_ptr()
            keep.append(table)
            jobs.append(jb)
        return jobs, keep

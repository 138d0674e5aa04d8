"""Per-trace output diagnostics (FP64 oracle) -- test infrastructure only
(see oracle/__init__.py).

Restates the reference's host-side O(T F) products:

* ``nadir``: the facet whose centre is horizontally nearest the platform
  (``output.py:69-72``; per interface in ``simulate.py:398-412``), ties to
  the lowest index (numpy argmin);
* ``first_return``: the earliest exact arrival 2|p - c|/c inside
  [t0, t0 + n dt) (``output.py:77-82``), out-of-window facets at +inf, so a
  trace with none reports facet 0 at +inf (argmin of all-inf);
* ``nadir_twtt_layers``: the vertical two-way time through the stack with
  in-medium speeds via each interface's nadir facet (``simulate.py:398-412``).

Evaluated trace by trace (the reference broadcasts (T, F) at once).
"""

import numpy as np


def nadir(positions, centers):
    """(index (T,), z (T,)) of the horizontally nearest facet centre."""
    pos = np.asarray(positions, np.float64)
    c = np.asarray(centers, np.float64)
    idx = np.empty(len(pos), np.int64)
    for t, p in enumerate(pos):
        d2 = ((c[:, :2] - p[None, :2]) ** 2).sum(-1)
        idx[t] = d2.argmin()
    return idx, c[idx, 2]


def first_return(positions, centers, t0, dt, n_samples, c):
    """(index (T,), twtt (T,), centre (T, 3)) of the earliest in-window
    facet arrival."""
    pos = np.asarray(positions, np.float64)
    cen = np.asarray(centers, np.float64)
    idx = np.empty(len(pos), np.int64)
    tw = np.empty(len(pos))
    for t, p in enumerate(pos):
        r = np.linalg.norm(cen - p[None], axis=-1)
        ft = 2.0 * r / c
        ft = np.where((ft >= t0) & (ft < t0 + n_samples * dt), ft, np.inf)
        idx[t] = ft.argmin()
        tw[t] = ft[idx[t]]
    return idx, tw, cen[idx]


def nadir_twtt_layers(positions, interfaces, eps, c):
    """(T, L) two-way time to each interface at nadir."""
    pos = np.asarray(positions, np.float64)
    out = np.empty((len(pos), len(interfaces)))
    prev = pos[:, 2].copy()
    acc = np.zeros(len(pos))
    for j, f in enumerate(interfaces):
        z = nadir(pos, f.centers)[1]
        acc = acc + np.sqrt(eps[j]) * (prev - z)
        out[:, j] = 2.0 * acc / c
        prev = z
    return out

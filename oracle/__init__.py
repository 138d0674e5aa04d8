"""FP64 CPU oracle for the coherent facet kernel -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import or
run it.  The library in ``radar_sounder_simulator_amd`` never imports it and
has no CPU fallback.

It restates, in float64/complex128 throughout (no float32 casts), the
reference's hot-path algorithm (soundersim, englacial/radar-sounder-simulator):

* ``coherent``   -- ``kernels/coherent.py:94-168, 193-244, 284-362`` (LPA
  facet field, rough variant, taper, antenna weighting, binning, drop-not-wrap,
  split sides, interp bins, reference-range phase fold-back);
* ``multilayer`` -- ``kernels/multilayer.py:187-591`` (component-form two-point
  Snell, sequential crossing chain with grid lookup, radiometry, binning);
* ``joint``      -- ``refraction_joint.py:109-322`` (joint N-interface Newton);
* ``roughness``  -- ``roughness.py:55-233``;
* ``antenna``    -- ``antenna.py:138-258`` (in-kernel gain functions);
* ``physics``    -- ``physics.py:16-67, 109-150``;
* ``waveform``   -- ``waveform.py:76-160``;
* ``scene``      -- ``compare/haynes.py:94-112`` (Gaussian-ACF height field)
  and the keyed white noise the device generator draws;
* ``brute``      -- referees ``compare/brute_force.py``, ``compare/
  brute_force_layered.py`` (used to PIN the oracle, not as the oracle);
* ``csrc/oracle.c`` -- the same surface and sequential-refracted math in C
  (OpenMP), used as the timed CPU baseline and cross-checked bit-exactly
  against the NumPy restatement.

Evaluation order is written out explicitly (dot products as ((x + y) + z),
products left to right) so that a float64 implementation that follows the
same order without FMA contraction reproduces bin indices and validity masks
bit for bit.

Pinning: every function here is checked by ``tests/test_oracle_*.py`` against
closed forms, brute-force referees, independent library routines (SciPy
``wofz``, mpmath, ``np.convolve``) or invariants -- see the per-function
docstrings and each module's header, which names its pins.
"""

"""Frozen algorithmic flop counts per (trace, facet) pair and MI355X peaks.

Counting rule: add, sub, mul, div, sqrt, floor and each transcendental (sin,
cos, exp, log, pow, atan2, acos) count 1; compares/selects/clamps count 0.
Counts follow the reference algorithm (its fixed Newton budgets 10 + 25 per
crossing, its per-pair evaluation of every factor), not what a particular
kernel executes, so both precision tiers are priced against the same work.

Surface pair (kernels/coherent.py:117-128, 219-236): 59 flops smooth and
isotropic -- d (3), r^2 (5), sqrt (1), d.n (5), /r (1), rhat (3), two sinc
terms 2 x (dot 5 + scale 1 + pi-mul 1 + sin 1 + div 1), amp (7), phase (2),
sin/cos (2), j amp e^{j phi} (2), twtt (2), bin (3), dropped power (3),
complex add (2).

Refracted crossing (kernels/multilayer.py:196-266, 371-404): two-point
solve = setup 56 + 29 per Newton iteration + epilogue 32; mean plane (10
iterations) 378 + lookup/renormalise 20 + local plane (25) 813 + cos 1 +
TE Fresnel 14 + transmission 3 + path sums 10 + cos 1 = 1240.
Final leg + spreading (:500-516) 42; antenna at departure 14 + gain;
coherent amplitude (:527-532) 27; phase + complex multiply 10; bin 5.

Focused SAR term (processing.py:318-331), one (output trace, aperture
trace, bin): 30 -- d^2 + r0^2 (2), sqrt (1), dr (1), tau (3), x (2), floor
and fraction (2), two complex linear interpolations (10), phase (1),
cos and sin (2), window product (2), complex multiply-add (4).

Incoherent surface pair (kernels/incoherent.py:36-53): 30 -- d (3), r^2
(5), sqrt (1), d.n (5), /r (1), A cos (1), its square (1), r^4 (2),
divide (1), twtt (2), bin (3), side (5); a pattern adds the departure
direction (3), the gain and g^4 (3).
"""

PEAK_FP32_TFLOPS = 157.3   # MI355X vector/matrix FP32, dense
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64, dense
PEAK_HBM_GBS = 8000.0

SURFACE_BASE = 59
FOCUS_TERM_FLOPS = 30
INCOHERENT_BASE = 30
CROSSING = 1240
FINAL_LEG = 42
REFR_AMP = 27
REFR_PHASE_BIN = 15
GAIN = {"isotropic": 0, "dipole": 8, "finite_dipole": 12, "array": 16,
        "array_tapered": 40, "tabulated": 10}
TAPER = 9
ROUGH_SETUP = 37          # l1, l2, K, A0, B0, x, log, amp_i, mean attenuation
ROUGH_PER_TERM_AREA = 16  # area-only term (Gaussian or exponential W_m)
ROUGH_PER_TERM_FULL = 260  # finite-facet Gaussian term: 4 Faddeeva (32-term
                           # complex Horner ~ 32 x 8) dominate


def rough_flops(n_terms, area_only=True):
    per = ROUGH_PER_TERM_AREA if area_only else ROUGH_PER_TERM_FULL
    return ROUGH_SETUP + per * int(n_terms) + 4


def surface_pair_flops(pattern="isotropic", taper=False, n_terms=0,
                       area_only=True):
    f = SURFACE_BASE + GAIN[pattern]
    f += TAPER if taper else 0
    f += rough_flops(n_terms, area_only) if n_terms else 0
    return f


def refracted_pair_flops(n_crossed=1, pattern="isotropic", taper=False,
                         n_terms=0, area_only=True):
    f = CROSSING * n_crossed + FINAL_LEG + REFR_AMP + REFR_PHASE_BIN
    f += (14 + GAIN[pattern]) if pattern != "isotropic" else 0
    f += TAPER if taper else 0
    f += rough_flops(n_terms, area_only) if n_terms else 0
    return f


def incoherent_pair_flops(pattern="isotropic"):
    return INCOHERENT_BASE + ((6 + GAIN[pattern]) if pattern != "isotropic" else 0)


# Bytes of a mixed refracted launch that no schedule can avoid: each facet
# record once (float64 centre + float32 normal, area, edges, gamma,
# phasors: 104 B), each crossed-interface lookup cell once (float64 centre
# and normal: 48 B), each trace record once (13 float64) and the complex128
# field once.
MIXED_FACET_BYTES = 104
TABLE_CELL_BYTES = 48
TRACE_RECORD_BYTES = 13 * 8


def refracted_unique_bytes(n_facets, n_traces, n_samples, table_cells=None):
    cells = n_facets if table_cells is None else table_cells
    return (MIXED_FACET_BYTES * n_facets + TABLE_CELL_BYTES * cells
            + TRACE_RECORD_BYTES * n_traces + 16 * n_traces * n_samples)

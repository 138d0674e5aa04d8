"""Sub-facet roughness (FP64), restating ``roughness.py:55-233``.

Gerekos et al. 2023 (Radio Science 58, doi:10.1029/2022RS007594): coherent
attenuation exp(-sigma^2 K^2/2) (Eq 20) and the incoherent variance D_Phi
(Eq 21) as a Poisson-weighted series, evaluated in log space.  The finite-
facet Gaussian terms use the Faddeeva function through the Weideman (1994)
N-term rational approximation; the area-only terms are the per-facet
infinite-surface law (Gaussian or exponential ACF spectrum W_m).

Pinned by tests/test_oracle_roughness.py: ``faddeeva`` vs scipy.special.wofz;
``d_phi`` vs an mpmath 50-digit evaluation of Eqs 21-24 and vs quadrature of
the area law; sigma = 0 exact zero; ``n_terms_for`` values; exponential W_m
vs direct numerical Hankel transform of exp(-r/l)^m; keyed phasors vs a
scalar pure-integer splitmix64.
"""

import math

import numpy as np

_INV_SQRT_PI = 1.0 / np.sqrt(np.pi)
_SQRT_PI = np.sqrt(np.pi)


def weideman_coeffs(n=32):
    """(L, a) constants of Weideman's rational approximation, highest power
    first (``roughness.py:55-64``)."""
    m = 2 * n
    k = np.arange(-m + 1, m)
    big_l = np.sqrt(n / np.sqrt(2.0))
    t = big_l * np.tan(np.pi * k / (2 * m))
    f = np.concatenate([[0.0], np.exp(-t * t) * (big_l * big_l + t * t)])
    a = np.real(np.fft.fft(np.fft.fftshift(f))) / (2 * m)
    return big_l, a[1:n + 1][::-1].copy()


def faddeeva(z, n=32):
    """w(z) = exp(-z^2) erfc(-iz) for Im z >= 0 (``roughness.py:67-80``)."""
    big_l, a = weideman_coeffs(n)
    z = np.asarray(z, np.complex128)
    iz = 1j * z
    d = big_l - iz
    zz = (big_l + iz) / d
    p = np.zeros_like(zz)
    for coef in a:              # Horner, highest power first (np.polyval)
        p = p * zz + coef
    return 2.0 * p / (d * d) + _INV_SQRT_PI / d


def mean_attenuation(sigma, K):
    """exp(-sigma^2 K^2 / 2) (``roughness.py:83-89``)."""
    sk = sigma * np.asarray(K, np.float64)
    return np.exp(-0.5 * (sk * sk))


def _f_factor(m, a0, edge, l):
    """One F factor of Eqs 22-24 via scaled Faddeeva (``roughness.py:92-107``)."""
    sq = np.sqrt(m)
    x = a0 * l / (2.0 * sq)
    y = edge * sq / l
    ex2 = np.exp(-x * x)
    ey2 = np.exp(-y * y)
    w_z = faddeeva(x + 1j * y)
    w_x = faddeeva(x + 0j)
    rot = np.exp(1j * (a0 * edge))
    t1 = -y * ex2 - ey2 * np.real(1j * (x + 1j * y) * rot * w_z)
    t2 = x * np.imag(w_x)
    return 1.0 - ey2 * np.cos(edge * a0) + _SQRT_PI * (t1 - t2)


def acf_spectrum(m, kb, l, acf="gaussian"):
    """W_m(k_B) of rho^m (``roughness.py:113-124``)."""
    if acf == "gaussian":
        return np.pi * (l * l) / m * np.exp(-(kb * kb) * (l * l) / (4.0 * m))
    if acf == "exponential":
        q = (l / m) ** 2
        return 2.0 * np.pi * q * (1.0 + (kb * kb) * q) ** -1.5
    raise ValueError(f"unknown roughness acf {acf!r}")


def d_phi(sigma, l, K, A0, B0, Lx, Ly, *, n_terms, area_only=False,
          acf="gaussian"):
    """Incoherent variance D_Phi (Eq 21), clamped >= 0 (``roughness.py:127-186``).

    log-space Poisson weight exp(m log x - lgamma(m+1) - x), x = (sigma K)^2;
    exactly 0 at x = 0."""
    if acf == "exponential" and not area_only:
        raise ValueError("acf='exponential' needs the area-only D_Phi")
    K, A0, B0, Lx, Ly = (np.asarray(v, np.float64) for v in (K, A0, B0, Lx, Ly))
    x = (sigma * K) ** 2
    with np.errstate(divide="ignore"):
        logx = np.where(x > 0, np.log(np.where(x > 0, x, 1.0)), -np.inf)
    acc = np.zeros(np.broadcast(x, A0, B0, Lx, Ly).shape)
    for m in range(1, n_terms + 1):
        mf = float(m)
        with np.errstate(invalid="ignore"):
            logp = np.where(x > 0, mf * logx, -np.inf) - math.lgamma(mf + 1.0) - x
        if area_only and acf == "exponential":
            kb = np.sqrt(A0 * A0 + B0 * B0)
            term = np.exp(logp) * acf_spectrum(mf, kb, l, "exponential") * Lx * Ly
        elif area_only:
            term = (np.exp(logp - (A0 * A0 + B0 * B0) * (l * l) / (4.0 * mf))
                    * (np.pi * (l * l) / mf) * Lx * Ly)
        else:
            l4 = (l * l) * (l * l)
            term = (np.exp(logp) * (l4 / (mf * mf))
                    * _f_factor(mf, A0, Lx, l) * _f_factor(mf, B0, Ly, l))
        acc = acc + term
    return np.maximum(acc, 0.0)


def n_terms_for(x_max):
    """Series length (``roughness.py:189-196``)."""
    x = float(x_max)
    return int(np.clip(np.ceil(x + 6.0 * np.sqrt(x) + 5.0), 10, 300))


def speckle_phasors(n, seed):
    """(eps1 + i eps2)/sqrt(2) from default_rng(seed) (``roughness.py:199-204``),
    complex128 here."""
    rng = np.random.default_rng(seed)
    ph = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return ph / np.sqrt(2.0)


def _splitmix_scalar(x):
    """splitmix64 finaliser on a Python int (mod 2^64)."""
    mask = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & mask
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & mask
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & mask
    return x ^ (x >> 31)


def speckle_phasors_for_keys(keys, seed):
    """Keyed Box-Muller phasors (``roughness.py:207-233``), scalar Python
    integer arithmetic per key (slow; small n only), complex128."""
    keys = np.asarray(keys, np.int64)
    base = int(np.random.SeedSequence(seed).generate_state(1, dtype=np.uint64)[0])
    mask = (1 << 64) - 1
    out = np.empty(len(keys), np.complex128)
    for i, (k0, k1) in enumerate(keys):
        keyed = base ^ (int(k0) & mask) ^ (((int(k1) & mask) * 0xD2B74407B1CE6E93) & mask)
        h1 = _splitmix_scalar(keyed)
        h2 = _splitmix_scalar(keyed ^ 0xCA5A826395121157)
        u1 = (float(h1 >> 11) + 0.5) * (1.0 / (1 << 53))
        u2 = (float(h2 >> 11) + 0.5) * (1.0 / (1 << 53))
        radius = math.sqrt(-2.0 * math.log(u1))
        angle = 2.0 * math.pi * u2
        out[i] = radius * (math.cos(angle) + 1j * math.sin(angle)) / math.sqrt(2.0)
    return out

// ss_math.h -- per-pair math of the coherent facet kernel, shared by the
// gfx950 kernels and by the CPU harness in tests/native (which compiles the
// very same functions with g++ to check them against the oracle).
//
// Two precision tiers:
//   *_fp64   float64 throughout, evaluation order written out and FMA
//            contraction disabled, so that bins, validity and crossing
//            geometry reproduce the FP64 oracle bit for bit.
//   *_mixed  float64 only where the error budget needs it (facet-platform
//            range, bin position, carrier-phase cycles, Snell crossing
//            polish); float32 for amplitudes, sinc, antenna, roughness.  A
//            pair whose bin position or lookup cell lies within a guard band
//            of a rounding boundary is recomputed on the float64 path, so
//            bins and masks stay identical to the float64 result.
#pragma once

#include <math.h>
#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SS_HD __host__ __device__ inline
#else
#define SS_HD inline
#endif

#include "../../include/soundersim_hip.h"

namespace ss {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kInvSqrtPi = 0.56418958354775628695;
constexpr double kSqrtPi = 1.77245385090551602730;
constexpr double kTiny = 1e-30;
constexpr double kEps64 = 2.220446049250313e-16;
constexpr double kCMin = 1e-9;
constexpr float kLog2E = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;

// ---------------------------------------------------------------------------
// float32 hardware primitives (gfx950: v_sin/v_cos take revolutions,
// v_rcp/v_rsq/v_sqrt/v_exp/v_log are the ~1 ulp transcendental units).
// The host harness uses libm so the same code runs on the CPU.
// ---------------------------------------------------------------------------
SS_HD float f_sin_rev(float x) {  // sin(2 pi x), |x| <= 0.5
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sinf(x);
#else
  return sinf((float)kTwoPi * x);
#endif
}
SS_HD float f_cos_rev(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_cosf(x);
#else
  return cosf((float)kTwoPi * x);
#endif
}
SS_HD float f_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
SS_HD float f_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
SS_HD float f_rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}
SS_HD float f_exp2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_exp2f(x);
#else
  return exp2f(x);
#endif
}
SS_HD float f_log2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_logf(x);
#else
  return log2f(x);
#endif
}
SS_HD float f_exp(float x) { return f_exp2(x * kLog2E); }
SS_HD float f_log(float x) { return f_log2(x) * kLn2; }

// sin(2 pi x) for any float x: exact reduction by the nearest integer.
SS_HD float f_sin_rev_any(float x) { return f_sin_rev(x - rintf(x)); }
SS_HD float f_cos_rev_any(float x) { return f_cos_rev(x - rintf(x)); }

// normalised sinc sin(pi x)/(pi x) in float32.
SS_HD float f_sinc(float x) {
  float ax = fabsf(x);
  if (ax < 1e-4f) {
    float y = (float)kPi * x;
    return 1.0f - y * y * (1.0f / 6.0f);
  }
  float h = 0.5f * x;
  float s = f_sin_rev(h - rintf(h));  // sin(pi x)
  return s * f_rcp((float)kPi * x);
}

// np.sinc in float64: y = pi * (x == 0 ? 1e-20 : x); sin(y)/y.
SS_HD double d_sinc(double x) {
#pragma clang fp contract(off)
  double y = kPi * (x == 0.0 ? 1.0e-20 : x);
  return sin(y) / y;
}

// ---------------------------------------------------------------------------
// Faddeeva w(z), Im z >= 0 (Weideman 1994, N = 32 coefficients a[] highest
// power first, constants table layout SS_CONST_*), and D_Phi (Gerekos 2023
// Eq 21) in either precision.
// ---------------------------------------------------------------------------
template <typename T>
struct cplx {
  T re, im;
};
template <typename T>
SS_HD cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
SS_HD cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  T den = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}

template <typename T>
SS_HD cplx<T> faddeeva(T x, T y, const double* cst) {
  const T L = (T)cst[SS_CONST_WEIDEMAN_L];
  // iz = i z = -y + i x; d = L - iz; zz = (L + iz)/d
  cplx<T> d = {L + y, -x};
  cplx<T> num = {L - y, x};
  cplx<T> zz = cdiv(num, d);
  cplx<T> p = {(T)0, (T)0};
  for (int i = 0; i < 32; ++i) {
    p = cmul(p, zz);
    p.re += (T)cst[SS_CONST_WEIDEMAN_A + i];
  }
  cplx<T> dd = cmul(d, d);
  cplx<T> t1 = cdiv(cplx<T>{(T)2 * p.re, (T)2 * p.im}, dd);
  cplx<T> t2 = cdiv(cplx<T>{(T)kInvSqrtPi, (T)0}, d);
  return {t1.re + t2.re, t1.im + t2.im};
}

template <typename T>
SS_HD T t_exp(T x) { return exp(x); }
template <>
SS_HD float t_exp<float>(float x) { return f_exp(x); }
template <typename T>
SS_HD T t_sqrt(T x) { return sqrt(x); }
template <>
SS_HD float t_sqrt<float>(float x) { return f_sqrt(x); }
template <typename T>
SS_HD T t_log(T x) { return log(x); }
template <>
SS_HD float t_log<float>(float x) { return f_log(x); }
template <typename T>
SS_HD T t_cos(T x) { return cos(x); }
template <>
SS_HD float t_cos<float>(float x) { return f_cos_rev_any(x * (float)(1.0 / kTwoPi)); }
template <typename T>
SS_HD T t_sin(T x) { return sin(x); }
template <>
SS_HD float t_sin<float>(float x) { return f_sin_rev_any(x * (float)(1.0 / kTwoPi)); }

// One F factor of Eqs 22-24 through the scaled Faddeeva combination.
template <typename T>
SS_HD T f_factor(T m, T a0, T edge, T l, const double* cst) {
  T sq = t_sqrt<T>(m);
  T x = a0 * l / ((T)2 * sq);
  T y = edge * sq / l;
  T ex2 = t_exp<T>(-x * x);
  T ey2 = t_exp<T>(-y * y);
  cplx<T> wz = faddeeva<T>(x, y, cst);
  cplx<T> wx = faddeeva<T>(x, (T)0, cst);
  T ang = a0 * edge;
  cplx<T> rot = {t_cos<T>(ang), t_sin<T>(ang)};
  // real(1j * (x + 1j y) * rot * w_z)
  cplx<T> u = {-y, x};
  cplx<T> v = cmul(cmul(u, rot), wz);
  T t1 = -y * ex2 - ey2 * v.re;
  T t2 = x * wx.im;
  return (T)1 - ey2 * t_cos<T>(ang) + (T)kSqrtPi * (t1 - t2);
}

// Per-term constants of the float32 area-only D_Phi series, staged per
// workgroup in LDS: entry m (m = 1 .. SS_MAX_TERMS) holds 1/m and m^2 as
// floats, read at a wave-uniform index (one broadcast LDS read per term).
constexpr int kTermTab = SS_MAX_TERMS + 1;
SS_HD void term_tab_entry(int m, float* tab) {
  tab[2 * m] = m > 0 ? (float)(1.0 / (double)m) : 0.0f;
  tab[2 * m + 1] = (float)((double)m * (double)m);
}

// D_Phi, clamped at 0 (exactly 0 when (sigma K)^2 == 0).
// RM (compile-time roughness mode): 1 area-only terms, 2 finite-facet
// Gaussian terms, -1 decided by ``area_only`` at run time.  ``tab``: the
// staged term constants (term_tab_entry) or null (computed per term);
// TAB_ALWAYS: the caller always passes the staged table (the float32 branch
// then carries no run-time choice between the two).
template <typename T, int RM = -1, bool TAB_ALWAYS = false>
SS_HD T d_phi(T sigma, T l, T K, T A0, T B0, T Lx, T Ly, int n_terms,
              bool area_only_rt, int acf, const double* cst, const float* tab = nullptr) {
  const bool area_only = RM == 1 ? true : (RM == 2 ? false : area_only_rt);
  T sk = sigma * K;
  T x = sk * sk;
  if (!(x > (T)0)) return (T)0;
  T acc = (T)0;
  T ab2 = A0 * A0 + B0 * B0;
  if constexpr (std::is_same<T, float>::value && RM != 2) {
    // float32 tier, area-only ACFs, e^-x a normal float (x < 80): the sum
    // over the Poisson weights P_m = e^-x x^m / m! in Horner form, from the
    // last term down, so no term needs its own weight, exp or lgamma:
    //   exponential, term P_m m (m^2 + c)^-3/2 with m P_m = x P_(m-1):
    //     x e^-x sum_(j<N) x^j/j! r_(j+1)^3,  G <- r_m^3 + (x/m) G
    //   Gaussian, term P_m e^(-c/4m) / m:
    //     e^-x sum_(m<=N) x^m/m! g_m,          G <- (x/m) (g_m + G)
    // with c = (A0^2 + B0^2) l^2 and the common factor l^2 applied once.
    // All terms are positive and G stays below e^x (r <= 1, g <= 1): no
    // cancellation, no overflow.  The log form (below) otherwise.
    if (area_only && x < 80.0f) {
      const float ll = l * l;
      const float c = ab2 * ll;
      const bool expo = acf == SS_ACF_EXPONENTIAL;
      const float cg = -0.25f * kLog2E * c;
      // one loop per (table, ACF) pair: both are uniform, and a branch on
      // either inside the loop costs more than the term
      auto horner = [&](auto tab_tag, auto expo_tag) -> float {
        constexpr bool TAB = decltype(tab_tag)::value;
        constexpr bool EXPO = decltype(expo_tag)::value;
        float G = 0.0f;
        float mf = (float)n_terms;
#pragma unroll 2
        for (int m = n_terms; m >= 1; --m) {
          float im, m2;
          if constexpr (TAB) {
            im = tab[2 * m];
            m2 = tab[2 * m + 1];
          } else {
            im = f_rcp(mf);
            m2 = mf * mf;
            mf -= 1.0f;
          }
          const float xm = x * im;
          if constexpr (EXPO) {
            const float r = f_rsq(m2 + c);
            G = xm * G + r * (r * r);
          } else {
            G = xm * (im * f_exp2(cg * im) + G);
          }
        }
        return G;
      };
      using TT_ = std::true_type;
      using FF_ = std::false_type;
      const float G = (TAB_ALWAYS || tab)
                          ? (expo ? horner(TT_{}, TT_{}) : horner(TT_{}, FF_{}))
                          : (expo ? horner(FF_{}, TT_{}) : horner(FF_{}, FF_{}));
      const float sum = (expo ? x : 1.0f) * f_exp(-x) * G;
      acc = sum * ll * ((expo ? (float)kTwoPi : (float)kPi) * Lx * Ly);
      return acc > (T)0 ? acc : (T)0;
    }
  }
  T logx = t_log<T>(x);
  if constexpr (std::is_same<T, double>::value && RM != 2) {
    // float64 tier, area-only exponential ACF: the Poisson weights by the
    // recurrence P_m = P_(m-1) x / m while e^-x is a normal double (x <
    // 700), no exp or lgamma per term (relative drift <= m ulp against
    // the log form, far inside the tier's 1e-9 parity)
    if (area_only && acf == SS_ACF_EXPONENTIAL && x < 700.0) {
      double P = exp(-x);
      double sum = 0.0;
      for (int m = 1; m <= n_terms; ++m) {
        const double mf = (double)m;
        P = P * x / mf;
        const double q = (l / mf) * (l / mf);
        const double rb = 1.0 / sqrt(1.0 + ab2 * q);
        sum += P * (q * (rb * rb * rb));
      }
      acc = sum * (kTwoPi * Lx * Ly);
      return acc > 0.0 ? acc : 0.0;
    }
  }
  for (int m = 1; m <= n_terms; ++m) {
    T mf = (T)m;
    T logp = mf * logx - (T)cst[SS_CONST_LGAMMA + m] - x;
    T term;
    if (RM != 2 && area_only && acf == SS_ACF_EXPONENTIAL) {
      T kb2 = ab2;
      T q = (l / mf) * (l / mf);
      T base = (T)1 + kb2 * q;
      // base^-1.5
      T rb = (T)1 / t_sqrt<T>(base);
      term = t_exp<T>(logp) * ((T)kTwoPi * q * (rb * rb * rb)) * Lx * Ly;
    } else if (RM != 2 && area_only) {
      term = t_exp<T>(logp - ab2 * (l * l) / ((T)4 * mf)) *
             ((T)kPi * (l * l) / mf) * Lx * Ly;
    } else if (RM != 1) {
      T l4 = (l * l) * (l * l);
      term = t_exp<T>(logp) * (l4 / (mf * mf)) * f_factor<T>(mf, A0, Lx, l, cst) *
             f_factor<T>(mf, B0, Ly, l, cst);
    } else {
      term = (T)0;
    }
    acc += term;
  }
  return acc > (T)0 ? acc : (T)0;
}

// ---------------------------------------------------------------------------
// Antenna: one-way FIELD gain at the departure direction (antenna.py:204-258)
// ---------------------------------------------------------------------------
struct PatternArgs {
  int kind;
  int n;               // tabulated samples / tapered elements
  double a;            // n_elements / length_lam
  double b[4];         // spacing_lam, field_scale, q, enabled
  const double* table; // tabulated: theta[n] | gain[n]; tapered: tx[n] | rx[n]
  // tabulated, staged lookup (pattern_lds_value / _index) in LDS, read by
  // antenna_gain<float, SS_PAT_TABULATED, true> (the incoherent kernel's
  // float32 tier); null elsewhere
  const float* ftab;   // theta[n] | gain[n] | slope[n - 1], float
  const short* fidx;   // kPatCells first-interval indices
  float fscale;        // kPatCells / (theta[n-1] - theta[0])
};

// Float32 tabulated-gain lookup staged per workgroup (patterns of at most
// kPatLds samples): the table rounded to float, the slope of every interval
// evaluated as the interpolation evaluates it, and for kPatCells uniform
// cells over [theta_0, theta_{n-1}] the interval holding each cell's lower
// edge, so a lookup starts within a step of its interval (finer cells than
// table intervals: one step up or down at most on a table whose samples are
// no closer than its range / kPatCells).
constexpr int kPatLds = 256;
constexpr int kPatCells = 1024;

SS_HD float pattern_cell_scale(const double* table, int n) {
  return (float)kPatCells / ((float)table[n - 1] - (float)table[0]);
}

// Entry i of the staged lookup (i < max(n, kPatCells)).
// Staged lookup of a tabulated pattern, in two passes over the workgroup
// (a barrier between them): pattern_lds_value(i) for i < n writes the
// float32 theta, gain and interval slope of sample i; pattern_lds_index(i)
// for i < kPatCells then writes the cell index -- the last interval whose
// start is <= the cell's start -- by a bisection over the staged float32
// thetas (the same comparisons as on (float) table values, read from LDS
// instead of a chain of dependent global loads).
SS_HD void pattern_lds_value(const double* table, int n, int i, float* ftab) {
  if (i < n) {
    ftab[i] = (float)table[i];
    ftab[n + i] = (float)table[n + i];
    if (i < n - 1) {
      const float x0 = (float)table[i], x1 = (float)table[i + 1];
      const float y0 = (float)table[n + i], y1 = (float)table[n + i + 1];
      ftab[2 * n + i] = (y1 - y0) / (x1 - x0);
    }
  }
}
SS_HD void pattern_lds_index(const float* ftab, int n, int i, short* fidx) {
  if (i < kPatCells) {
    const float xa = ftab[0];
    const float e = xa + (float)i / ((float)kPatCells / (ftab[n - 1] - xa));
    int lo = 0, hi = n - 1;  // last interval whose start is <= e
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ftab[mid] <= e)
        lo = mid;
      else
        hi = mid;
    }
    fidx[i] = (short)lo;
  }
}

// Linear interpolation in a staged tabulated pattern at theta = th: the
// interval xf[lo] <= th < xf[lo + 1] (the one a bisection finds) from the
// cell index and a step or two, then slope * (th - x_lo) + y_lo.
SS_HD float tab_lookup_staged(const PatternArgs& pa, float th) {
  const float* xf = pa.ftab;
  const float* ff = pa.ftab + pa.n;
  const float* sl = pa.ftab + 2 * pa.n;
  const int n = pa.n;
  const float xa = xf[0], xb = xf[n - 1];
  int cell = (int)((fminf(fmaxf(th, xa), xb) - xa) * pa.fscale);
  cell = cell < 0 ? 0 : (cell > kPatCells - 1 ? kPatCells - 1 : cell);
  int lo = pa.fidx[cell];
  // one step either way (the cell index rounds), then the general walk for
  // tables with several samples in a cell, rarely entered
  lo -= (lo > 0 && xf[lo] > th) ? 1 : 0;
  lo += (lo < n - 2 && xf[lo + 1] <= th) ? 1 : 0;
  if (lo < n - 2 && xf[lo + 1] <= th) {
    do ++lo;
    while (lo < n - 2 && xf[lo + 1] <= th);
  } else if (lo > 0 && xf[lo] > th) {
    do --lo;
    while (lo > 0 && xf[lo] > th);
  }
  float g = sl[lo] * (th - xf[lo]) + ff[lo];
  g = th < xb ? g : ff[n - 1];
  return th > xa ? g : ff[0];  // (NaN: ff[0], as the bisection's first test)
}

// acos in float32 after Abramowitz & Stegun (1964) 4.4.46 (Hastings):
// acos(x) = sqrt(1 - x) (a0 + a1 x + ... + a7 x^7) on [0, 1], |error| <=
// 2e-8 rad, and acos(-x) = pi - acos(x).
SS_HD float f_acos(float x) {
  const float ax = fabsf(x);
  float p = -0.0012624911f;
  p = p * ax + 0.0066700901f;
  p = p * ax - 0.0170881256f;
  p = p * ax + 0.0308918810f;
  p = p * ax - 0.0501743046f;
  p = p * ax + 0.0889789874f;
  p = p * ax - 0.2145988016f;
  p = p * ax + 1.5707963050f;
  const float r = f_sqrt(1.0f - ax) * p;
  return x < 0.0f ? (float)kPi - r : r;
}

template <typename T>
SS_HD T t_pow(T x, T y) { return pow(x, y); }
template <>
SS_HD float t_pow<float>(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x <= 0.0f) return 0.0f;
  return f_exp2(y * f_log2(x));
}
template <typename T>
SS_HD T t_acos(T x) { return acos(x); }

template <typename T>
SS_HD T element_field(T c, T q, bool enabled) {
  if (!enabled) return (T)1;
  if (!(c >= (T)0)) return (T)0;
  T cl = c > (T)1 ? (T)1 : c;
  return t_pow<T>(cl, q / (T)2);
}

// dh: departure unit vector; pv0/pv1: trace pattern vectors.
// PAT: compile-time pattern kind (SS_PAT_*), or -1 to dispatch on pa.kind.
// STAGED (float32, tabulated): the caller staged the table's lookup in
// ``pa.ftab`` / ``pa.fidx``; theta from f_acos.
template <typename T, int PAT = -1, bool STAGED = false>
SS_HD T antenna_gain(const PatternArgs& pa, const T dh[3], const T pv0[3],
                     const T pv1[3]) {
  switch (PAT >= 0 ? PAT : pa.kind) {
    case SS_PAT_DIPOLE: {
      T ca = (dh[0] * pv0[0] + dh[1] * pv0[1]) + dh[2] * pv0[2];
      T s2 = (T)1 - ca * ca;
      T s = t_sqrt<T>(s2 > (T)1e-12 ? s2 : (T)1e-12);
      return t_cos<T>((T)(kPi / 2.0) * ca) / s;
    }
    case SS_PAT_FINITE_DIPOLE: {
      T kh = (T)kPi * (T)pa.a;
      T ca = (dh[0] * pv0[0] + dh[1] * pv0[1]) + dh[2] * pv0[2];
      T s2 = (T)1 - ca * ca;
      T s = t_sqrt<T>(s2 > (T)1e-12 ? s2 : (T)1e-12);
      T ckh = t_cos<T>(kh);
      return fabs(t_cos<T>(kh * ca) - ckh) / (s * ((T)1 - ckh));
    }
    case SS_PAT_ARRAY: {
      T u = (dh[0] * pv0[0] + dh[1] * pv0[1]) + dh[2] * pv0[2];
      T ct = (dh[0] * pv1[0] + dh[1] * pv1[1]) + dh[2] * pv1[2];
      T na = (T)pa.a;
      T x = (T)kPi * (T)pa.b[0] * u;
      T sx = t_sin<T>(x);
      T af;
      if (fabs(sx) < (T)1e-5)
        af = t_cos<T>(na * x) / t_cos<T>(x);
      else
        af = t_sin<T>(na * x) / (na * sx);
      T ef = element_field<T>(ct, (T)pa.b[2], pa.b[3] != 0.0);
      return (T)pa.b[1] * ef * fabs(af);
    }
    case SS_PAT_ARRAY_TAPERED: {
      T u = (dh[0] * pv0[0] + dh[1] * pv0[1]) + dh[2] * pv0[2];
      T ct = (dh[0] * pv1[0] + dh[1] * pv1[1]) + dh[2] * pv1[2];
      int n = pa.n;
      T ctx = 0, stx = 0, crx = 0, srx = 0, wt = 0, wr = 0;
      T half = (T)(n - 1) / (T)2;
      for (int i = 0; i < n; ++i) {
        T mi = (T)i - half;
        T ph = ((T)kTwoPi * (T)pa.b[0]) * u * mi;
        T c = t_cos<T>(ph), s = t_sin<T>(ph);
        T w1 = (T)pa.table[i], w2 = (T)pa.table[n + i];
        ctx += c * w1;
        stx += s * w1;
        crx += c * w2;
        srx += s * w2;
        wt += w1;
        wr += w2;
      }
      T aft = t_sqrt<T>(ctx * ctx + stx * stx) / wt;
      T afr = t_sqrt<T>(crx * crx + srx * srx) / wr;
      T ef = element_field<T>(ct, (T)pa.b[2], pa.b[3] != 0.0);
      return (T)pa.b[1] * ef * t_sqrt<T>(aft * afr);
    }
    case SS_PAT_TABULATED: {
      T ct = (dh[0] * pv0[0] + dh[1] * pv0[1]) + dh[2] * pv0[2];
      ct = ct < (T)-1 ? (T)-1 : (ct > (T)1 ? (T)1 : ct);
      if constexpr (STAGED && std::is_same<T, float>::value)
        return tab_lookup_staged(pa, f_acos(ct));
      T th = t_acos<T>(ct);
      const double* xp = pa.table;
      const double* fp = pa.table + pa.n;
      int n = pa.n;
      if (!(th > (T)xp[0])) return (T)fp[0];
      if (!(th < (T)xp[n - 1])) return (T)fp[n - 1];
      int lo = 0, hi = n - 1;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if ((T)xp[mid] <= th)
          lo = mid;
        else
          hi = mid;
      }
      T x0 = (T)xp[lo], x1 = (T)xp[hi];
      T y0 = (T)fp[lo], y1 = (T)fp[hi];
      T slope = (y1 - y0) / (x1 - x0);
      return slope * (th - x0) + y0;
    }
    default:
      return (T)1;
  }
}

// ---------------------------------------------------------------------------
// Per-trace and per-facet records
// ---------------------------------------------------------------------------
struct TraceRec {
  double p[3];
  double u[3];
  double rref;
  double pv0[3];
  double pv1[3];
};

struct SurfaceConst {
  double k, gamma, t0, dt, c, taper_s, sigma, lc;
  int n_samples, flags, n_terms, acf;
  PatternArgs pat;
  const double* cst;
  const float* mtab = nullptr;   // staged D_Phi term constants (LDS) or null
  // derived (host-filled)
  double k_over_pi, kfac;        // k/pi, k/(2 pi)
  double k_bin, t0_dt;           // 2/(c dt), t0/dt   (mixed fast bin)
  double cyc_per_m;              // 2/lambda = k/pi   (mixed phase cycles)
};

struct SurfFacet64 {
  double c[3], n[3], A, e1[3], e2[3], gam, ph_re, ph_im, l1, l2;
};

struct SurfFacet32 {
  double c[3];  // float64 centre (range and bin need it)
  float n[3], A, e1[3], e2[3], gam, ph_re, ph_im, l1, l2;
};

// Result of one pair: up to two binned parts (interp bins) + side.
struct PairOut {
  double re[2], im[2];
  int bin[2];      // floor bins (part 1 = bin + 1 with interp)
  int nparts;
  int right;       // split sides: 1 when the facet is right of track
  int raw_bin;     // diagnostic: floor bin of the pair (or SS_BIN_INVALID)
  bool valid;      // path validity (refracted); always true for the surface
  bool redo;       // reduced-precision path unresolved: the float64 redo
                   // kernel owns this pair (it contributes nothing here)
  bool power;      // re holds a power (incoherent): an out-of-window part
                   // books re itself to dropped, not |re + j im|^2
  double pwr_invalid;  // |contrib|^2 when the path is invalid
};

SS_HD double exact_bin_pos(double r_or_opl, double c, double t0, double dt) {
#pragma clang fp contract(off)
  double twtt = 2.0 * r_or_opl / c;
  return (twtt - t0) / dt;
}

// ---------------------------------------------------------------------------
// Surface pair, float64 (kernels/coherent.py:107-168, 216-236)
// ---------------------------------------------------------------------------
template <int RM = -1, int PAT = -1>
SS_HD void surface_pair_fp64(const SurfFacet64& f, const TraceRec& t,
                             const SurfaceConst& sc, PairOut& o) {
#pragma clang fp contract(off)
  const int flags = sc.flags;
  const bool rough = RM != 0 && (flags & SS_ROUGH);
  const bool pattern = PAT != 0 && sc.pat.kind != SS_PAT_ISOTROPIC;
  double dx = t.p[0] - f.c[0], dy = t.p[1] - f.c[1], dz = t.p[2] - f.c[2];
  double r = sqrt((dx * dx + dy * dy) + dz * dz);
  double cs = ((dx * f.n[0] + dy * f.n[1]) + dz * f.n[2]) / r;
  double rx = dx / r, ry = dy / r, rz = dz / r;
  double d1 = (rx * f.e1[0] + ry * f.e1[1]) + rz * f.e1[2];
  double d2 = (rx * f.e2[0] + ry * f.e2[1]) + rz * f.e2[2];
  double s1 = d_sinc(d1 * (sc.k / kPi));
  double s2 = d_sinc(d2 * (sc.k / kPi));
  double gam = (flags & SS_GAMMA_FACET) ? f.gam : sc.gamma;
  double amp = (sc.k / kTwoPi) * gam * cs * f.A * s1 * s2 / (r * r);
  if (flags & SS_TAPER) {
    double cc = cs < 1e-9 ? 1e-9 : (cs > 1.0 ? 1.0 : cs);
    double c2 = cc * cc;
    amp = amp * exp(-(1.0 - c2) / (c2 * (2.0 * sc.taper_s * sc.taper_s)));
  }
  double phase = (-2.0 * sc.k) * (r - t.rref);
  double cre = amp, cim = 0.0;
  if (rough) {
    double kk = 2.0 * sc.k * cs;
    double dp = 0.0;
    if (f.A > 0.0)
      dp = d_phi<double, RM>(sc.sigma, sc.lc, kk, 2.0 * sc.k * d1 / f.l1,
                         2.0 * sc.k * d2 / f.l2, f.l1, f.l2, sc.n_terms,
                         (flags & SS_D_PHI_AREA) != 0, sc.acf, sc.cst);
    double amp_i = f.A > 0.0 ? (sc.k / kTwoPi) * gam * cs / (r * r) * sqrt(dp) : 0.0;
    double sk = sc.sigma * kk;
    double matt = exp(-0.5 * (sk * sk));
    cre = amp * matt + amp_i * f.ph_re;
    cim = amp_i * f.ph_im;
  }
  double sp = sin(phase), cp = cos(phase);
  // j (cre + j cim) (cp + j sp)
  double re = -(cre * sp + cim * cp);
  double im = cre * cp - cim * sp;
  if (pattern) {
    double dh[3] = {-dx / r, -dy / r, -dz / r};
    double g = antenna_gain<double, PAT>(sc.pat, dh, t.pv0, t.pv1);
    double g2 = g * g;
    re *= g2;
    im *= g2;
  }
  double twtt = 2.0 * r / sc.c;
  double pos = (twtt - sc.t0) / sc.dt;
  double fb = floor(pos);
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.raw_bin = b;
  o.valid = true;
  o.pwr_invalid = 0.0;
  if (flags & SS_INTERP_BINS) {
    double w = pos - fb;
    o.nparts = 2;
    o.re[0] = re * (1.0 - w);
    o.im[0] = im * (1.0 - w);
    o.bin[0] = b;
    o.re[1] = re * w;
    o.im[1] = im * w;
    o.bin[1] = b + 1;
  } else {
    o.nparts = 1;
    o.re[0] = re;
    o.im[0] = im;
    o.bin[0] = b;
  }
  o.right = (((-dx) * t.u[0] + (-dy) * t.u[1]) + (-dz) * t.u[2]) > 0.0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Surface pair, mixed precision
// ---------------------------------------------------------------------------
constexpr double kBinGuard = 1e-7;   // bin units; fast-path error is ~1e-10

template <int RM = -1, int PAT = -1, bool STAGED = false>
SS_HD void surface_pair_mixed(const SurfFacet32& f, const SurfFacet64* f64,
                              const TraceRec& t, const SurfaceConst& sc,
                              PairOut& o) {
  const int flags = sc.flags;
  const bool rough = RM != 0 && (flags & SS_ROUGH);
  const bool pattern = PAT != 0 && sc.pat.kind != SS_PAT_ISOTROPIC;
  double dx = t.p[0] - f.c[0], dy = t.p[1] - f.c[1], dz = t.p[2] - f.c[2];
  double r2 = dx * dx + dy * dy + dz * dz;
  float r0 = f_sqrt((float)r2);
  float ir0 = f_rcp(r0);
  double res = fma(-(double)r0, (double)r0, r2);
  float corr = (float)res * (0.5f * ir0);
  double r = (double)r0 + (double)corr;
  // bin position (fast) with exact fallback near a bin edge
  double pos = fma(r, sc.k_bin, -sc.t0_dt);
  double fb = floor(pos);
  double fr = pos - fb;
  if (fr < kBinGuard || fr > 1.0 - kBinGuard) {
    pos = exact_bin_pos(sqrt((dx * dx + dy * dy) + dz * dz), sc.c, sc.t0, sc.dt);
    fb = floor(pos);
  }
  // carrier phase in cycles: -2 k r = -2 pi (r * 2/lambda)
  double cyc = r * sc.cyc_per_m;
  float ph = (float)(cyc - rint(cyc));
  float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
  float cs = (dxf * f.n[0] + dyf * f.n[1] + dzf * f.n[2]) * ir0;
  float rx = dxf * ir0, ry = dyf * ir0, rz = dzf * ir0;
  float d1 = rx * f.e1[0] + ry * f.e1[1] + rz * f.e1[2];
  float d2 = rx * f.e2[0] + ry * f.e2[1] + rz * f.e2[2];
  float kop = (float)sc.k_over_pi;
  float s1 = f_sinc(d1 * kop);
  float s2 = f_sinc(d2 * kop);
  float gam = (flags & SS_GAMMA_FACET) ? f.gam : (float)sc.gamma;
  float kg = (float)sc.kfac * gam;
  float ir2 = ir0 * ir0;
  float amp = kg * cs * f.A * s1 * s2 * ir2;
  if (flags & SS_TAPER) {
    float cc = cs < 1e-9f ? 1e-9f : (cs > 1.0f ? 1.0f : cs);
    float c2 = cc * cc;
    float ts = (float)sc.taper_s;
    amp *= f_exp(-(1.0f - c2) * f_rcp(c2 * (2.0f * ts * ts)));
  }
  float cre = amp, cim = 0.0f;
  if (rough) {
    float k2 = 2.0f * (float)sc.k;
    float kk = k2 * cs;
    float dp = 0.0f;
    if (f.A > 0.0f)
      dp = d_phi<float, RM>((float)sc.sigma, (float)sc.lc, kk, k2 * d1 * f_rcp(f.l1),
                        k2 * d2 * f_rcp(f.l2), f.l1, f.l2, sc.n_terms,
                        (flags & SS_D_PHI_AREA) != 0, sc.acf, sc.cst, sc.mtab);
    float amp_i = f.A > 0.0f ? kg * cs * ir2 * f_sqrt(dp) : 0.0f;
    float sk = (float)sc.sigma * kk;
    float matt = f_exp(-0.5f * sk * sk);
    cre = amp * matt + amp_i * f.ph_re;
    cim = amp_i * f.ph_im;
  }
  // e^{j phase}, phase = -2 pi ph: cos = cos(2 pi ph), sin = -sin(2 pi ph)
  float sp = -f_sin_rev(ph), cp = f_cos_rev(ph);
  float re = -(cre * sp + cim * cp);
  float im = cre * cp - cim * sp;
  if (pattern) {
    float dh[3] = {-rx, -ry, -rz};
    float pv0[3] = {(float)t.pv0[0], (float)t.pv0[1], (float)t.pv0[2]};
    float pv1[3] = {(float)t.pv1[0], (float)t.pv1[1], (float)t.pv1[2]};
    float g = antenna_gain<float, PAT, STAGED>(sc.pat, dh, pv0, pv1);
    float g2 = g * g;
    re *= g2;
    im *= g2;
  }
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.raw_bin = b;
  o.valid = true;
  o.pwr_invalid = 0.0;
  if (flags & SS_INTERP_BINS) {
    float w = (float)(pos - fb);
    o.nparts = 2;
    o.re[0] = re * (1.0f - w);
    o.im[0] = im * (1.0f - w);
    o.bin[0] = b;
    o.re[1] = re * w;
    o.im[1] = im * w;
    o.bin[1] = b + 1;
  } else {
    o.nparts = 1;
    o.re[0] = re;
    o.im[0] = im;
    o.bin[0] = b;
  }
  o.right = (-(dxf * (float)t.u[0] + dyf * (float)t.u[1] + dzf * (float)t.u[2])) > 0.0f ? 1 : 0;
  (void)f64;
}

// ---------------------------------------------------------------------------
// Incoherent surface pair (kernels/incoherent.py:36-53): power
// (A cos)^2 / r^4, times g^4 for a non-isotropic antenna at the departure
// direction (c - p)/r; bin floor((2r/c - t0)/dt); side (c - p).u_ct > 0.
// The part's re is the power (o.power): an out-of-window part books it to
// dropped as it is.
// ---------------------------------------------------------------------------
struct IncohFacet {
  double c[3], n[3], A;
};

template <int PAT = -1>
SS_HD void incoherent_pair_fp64(const IncohFacet& f, const TraceRec& t, const SurfaceConst& sc,
                                PairOut& o) {
#pragma clang fp contract(off)
  double dx = t.p[0] - f.c[0], dy = t.p[1] - f.c[1], dz = t.p[2] - f.c[2];
  double r = sqrt((dx * dx + dy * dy) + dz * dz);
  double cs = ((dx * f.n[0] + dy * f.n[1]) + dz * f.n[2]) / r;
  double a = f.A * cs;
  double pw = (a * a) / ((r * r) * (r * r));
  if (PAT != 0 && sc.pat.kind != SS_PAT_ISOTROPIC) {
    double dh[3] = {-dx / r, -dy / r, -dz / r};
    double g = antenna_gain<double, PAT>(sc.pat, dh, t.pv0, t.pv1);
    pw = pw * ((g * g) * (g * g));
  }
  double fb = floor((2.0 * r / sc.c - sc.t0) / sc.dt);
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.raw_bin = b;
  o.valid = true;
  o.power = true;
  o.pwr_invalid = 0.0;
  o.nparts = 1;
  o.re[0] = pw;
  o.im[0] = 0.0;
  o.bin[0] = b;
  o.right = (((-dx) * t.u[0] + (-dy) * t.u[1]) + (-dz) * t.u[2]) > 0.0 ? 1 : 0;
}

// Mixed tier: the float64 range and bin of the mixed surface pair (float32
// square root with one float64 correction, exact fallback within 1e-7 bin
// of an edge, so bins equal the float64 tier's) and the float64 side,
// float32 power and gain.
template <int PAT = -1>
SS_HD void incoherent_pair_mixed(const IncohFacet& f, const TraceRec& t, const SurfaceConst& sc,
                                 PairOut& o) {
  double dx = t.p[0] - f.c[0], dy = t.p[1] - f.c[1], dz = t.p[2] - f.c[2];
  double r2 = dx * dx + dy * dy + dz * dz;
  float r0 = f_sqrt((float)r2);
  float ir0 = f_rcp(r0);
  double res = fma(-(double)r0, (double)r0, r2);
  float corr = (float)res * (0.5f * ir0);
  double r = (double)r0 + (double)corr;
  double pos = fma(r, sc.k_bin, -sc.t0_dt);
  double fb = floor(pos);
  double fr = pos - fb;
  if (fr < kBinGuard || fr > 1.0 - kBinGuard) {
    pos = exact_bin_pos(sqrt((dx * dx + dy * dy) + dz * dz), sc.c, sc.t0, sc.dt);
    fb = floor(pos);
  }
  float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
  float cs = (dxf * (float)f.n[0] + dyf * (float)f.n[1] + dzf * (float)f.n[2]) * ir0;
  float a = (float)f.A * cs;
  float ir2 = ir0 * ir0;
  float pw = (a * a) * (ir2 * ir2);
  if (PAT != 0 && sc.pat.kind != SS_PAT_ISOTROPIC) {
    float dh[3] = {-dxf * ir0, -dyf * ir0, -dzf * ir0};
    float pv0[3] = {(float)t.pv0[0], (float)t.pv0[1], (float)t.pv0[2]};
    float pv1[3] = {(float)t.pv1[0], (float)t.pv1[1], (float)t.pv1[2]};
    // the incoherent kernel stages a tabulated pattern (PAT == 5) in LDS
    float g = antenna_gain<float, PAT, PAT == SS_PAT_TABULATED>(sc.pat, dh, pv0, pv1);
    float g2 = g * g;
    pw *= g2 * g2;
  }
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.raw_bin = b;
  o.valid = true;
  o.power = true;
  o.pwr_invalid = 0.0;
  o.nparts = 1;
  o.re[0] = pw;
  o.im[0] = 0.0;
  o.bin[0] = b;
  // float64 side: facets straight below the track sit at a cross-track
  // offset near zero
  o.right = (((-dx) * t.u[0] + (-dy) * t.u[1]) + (-dz) * t.u[2]) > 0.0 ? 1 : 0;
}

}  // namespace ss

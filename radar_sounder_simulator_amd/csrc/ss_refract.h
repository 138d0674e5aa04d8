// ss_refract.h -- refracted-path pair math (sequential two-point chain).
//
// Float64 tier: a line-by-line float64 evaluation of the safeguarded
// sigma-Newton two-point solve (kernels/multilayer.py:196-266), the mean-plane
// -> cell lookup -> local-plane chain (:352-406) and the radiometry
// (:500-583), contraction off, so crossing points, validity, optical path and
// bins reproduce the FP64 oracle.
//
// Mixed tier: the mean-plane anchor is solved in float32 in plane-relative
// coordinates (its only output is a lookup cell; a pair whose cell
// coordinate is within kCellGuard of a rounding boundary redoes it in float64
// with the reference budget); the local-plane solve sets up in float64,
// iterates Newton in float32 and polishes with float64 Newton steps, so the
// crossing, residual and optical path reach float64 accuracy; the
// radiometric factors run in float32.  Pairs whose bin position is within
// kBinGuard of an edge, or whose Snell residual falls in the ambiguous band
// around the 1e-9 tolerance, are recomputed on the float64 tier.
#pragma once

#include "ss_math.h"

namespace ss {

constexpr int kMeanIters = 10;     // reference budget, mean-plane anchor
constexpr int kLocalIters = 25;    // reference budget, local plane
constexpr int kMeanIters32 = 12;   // mixed: float32 anchor iterations
constexpr int kLocalIters32 = 10;  // mixed: float32 local iterations
constexpr int kPolishMax = 3;      // mixed: float64 polish steps at most (one unless
                                   // the residual is ambiguous after it)

// Mixed tier, float64: 1/x and 1/sqrt(x) from the hardware estimates
// refined by two Newton-Raphson steps (r <- r (2 - x r); y <- y (3 - x y^2)
// / 2), accurate to about an ulp without the IEEE division's scale / fixup
// sequence.  The host build divides.
SS_HD double rcp64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
SS_HD double rsq64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
#else
  return 1.0 / sqrt(x);
#endif
}
// 1/sqrt(x) from the hardware estimate (about 2^-23 relative) refined by
// ONE Newton-Raphson step: about 2e-14 relative, for sums whose float64
// result needs far less than an ulp (the flat-stack layer loop: optical
// path and horizontal reach to 1e-10 m over kilometres).  The host build
// starts from a float32 estimate so the CPU tests see the same error class.
SS_HD double rsq64_1(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
#else
  double y = (double)(1.0f / sqrtf((float)x));
#endif
  const double e = fma(-(0.5 * x) * y, y, 0.5);
  return fma(y, e, y);
}
// sqrt as x rsq64(x), keeping sqrt(0) = 0 and sqrt(x < 0) = NaN.
SS_HD double sqrt64(double x) {
  const double y = x * rsq64(x > 0.0 ? x : 1.0);
  return x > 0.0 ? y : (x == 0.0 ? x : __builtin_nan(""));
}
constexpr double kCellGuard = 2e-4;  // cells; float32 anchor error ~1e-5
constexpr double kResLo = 1e-11, kResHi = 1e-7;  // ambiguous residual band
// float32 Newton loops stop once a full step moves sigma by at most this
// relative amount (4 ulp): settled, or alternating between neighbours
constexpr float kF32Settle = 4.8e-7f;
// Mixed local-plane solve, polish-free crossing: bound of the settled
// float32 crossing's in-plane error relative to the in-plane span L (about
// 1.5x the rounding estimate), and the sigma = sin(theta) below which it
// applies (grazing incidence always takes the float64 polish).
constexpr float kPosRelErr = 2e-6f;
constexpr float kSigFast = 0.995f;
// ... with its optical-path error at most this fraction of the bin guard
// and its legs' direction error at most kDirRelErr (radians)
constexpr double kFastOplFrac = 0.25;
constexpr float kDirRelErr = 1e-5f;

struct Snell64 {
  double x[3];
  double th1, th2, s1, s2, res;
  bool valid;
};

// refraction.py:97-161 / kernels/multilayer.py:196-266, float64.
SS_HD void snell_fp64(const double p[3], const double q[3], const double o[3],
                      const double nrm[3], double n1, double n2, int n_iter,
                      Snell64& out) {
#pragma clang fp contract(off)
  double h1 = ((p[0] - o[0]) * nrm[0] + (p[1] - o[1]) * nrm[1]) + (p[2] - o[2]) * nrm[2];
  double h2 = ((q[0] - o[0]) * nrm[0] + (q[1] - o[1]) * nrm[1]) + (q[2] - o[2]) * nrm[2];
  double a = fabs(h1), b = fabs(h2);
  double fp[3] = {p[0] - h1 * nrm[0], p[1] - h1 * nrm[1], p[2] - h1 * nrm[2]};
  double fq[3] = {q[0] - h2 * nrm[0], q[1] - h2 * nrm[1], q[2] - h2 * nrm[2]};
  double w[3] = {fq[0] - fp[0], fq[1] - fp[1], fq[2] - fp[2]};
  double L = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  double Lm = L > kTiny ? L : kTiny;
  double u[3] = {w[0] / Lm, w[1] / Lm, w[2] / Lm};
  double sm = (a + b) + L;
  double s = sm > kTiny ? sm : kTiny;
  double inv_s = 1.0 / s;
  double ah = a * inv_s, bh = b * inv_s, Lh = L * inv_s;
  bool swap = n1 > n2;
  double ratio = (n1 < n2 ? n1 : n2) / (n1 > n2 ? n1 : n2);
  double ar = swap ? bh : ah;
  double bd = swap ? ah : bh;
  double hi0 = 1.0 - 8.0 * kEps64;
  double ab = ah + bh;
  double den = sqrt(Lh * Lh + ab * ab);
  double sig = Lh / (den > kTiny ? den : kTiny);
  sig = sig < hi0 ? sig : hi0;
  double lo = 0.0, hi = hi0;
  for (int it = 0; it < n_iter; ++it) {
    double v = ratio * sig;
    double c1 = (1.0 - sig) * (1.0 + sig);
    c1 = c1 > kTiny ? c1 : kTiny;
    double c2 = (1.0 - v) * (1.0 + v);
    c2 = c2 > kTiny ? c2 : kTiny;
    double F = (ar * sig * (1.0 / sqrt(c1)) + bd * v * (1.0 / sqrt(c2))) - Lh;
    bool le = F <= 0.0;
    lo = le ? sig : lo;
    hi = le ? hi : sig;
    double Fp = ar / (c1 * sqrt(c1)) + (bd * ratio) / (c2 * sqrt(c2));
    double sn = sig - F / (Fp > kTiny ? Fp : kTiny);
    bool ok = (sn >= lo) && (sn <= hi) && isfinite(sn);
    sig = ok ? sn : 0.5 * (lo + hi);
  }
  double c1 = (1.0 - sig) * (1.0 + sig);
  c1 = c1 > kTiny ? c1 : kTiny;
  double tr = ar * sig * (1.0 / sqrt(c1));
  double tt = swap ? Lh - tr : tr;
  double t = tt < 0.0 ? 0.0 : (tt > Lh ? Lh : tt);
  double d2 = Lh - t;
  double r1 = sqrt(ah * ah + t * t);
  r1 = r1 > kTiny ? r1 : kTiny;
  double r2 = sqrt(bh * bh + d2 * d2);
  r2 = r2 > kTiny ? r2 : kTiny;
  out.res = n1 * t / r1 - n2 * d2 / r2;
  out.s1 = s * r1;
  out.s2 = s * r2;
  double st = s * t;
  out.x[0] = fp[0] + st * u[0];
  out.x[1] = fp[1] + st * u[1];
  out.x[2] = fp[2] + st * u[2];
  out.th1 = atan2(t, ah);
  out.th2 = atan2(d2, bh);
  out.valid = (h1 * h2 < 0.0) && (fabs(out.res) <= 1e-9);
}

// Mixed local-plane solve: float64 setup, float32 Newton, float64 polish.
// Returns the crossing, s1, incidence/refraction cosines and validity.
struct SnellMixed {
  double x[3];
  double s1;
  float c_in, c_out;  // cos(theta1), cos(theta2) (unclamped)
  double res;
  bool valid;
  bool ambiguous;  // residual inside the guard band around the tolerance
};

SS_HD void snell_mixed(const double p[3], const double q[3], const double o[3],
                       const double nrm[3], double n1, double n2, float opl_tol,
                       SnellMixed& out) {
  double h1 = ((p[0] - o[0]) * nrm[0] + (p[1] - o[1]) * nrm[1]) + (p[2] - o[2]) * nrm[2];
  double h2 = ((q[0] - o[0]) * nrm[0] + (q[1] - o[1]) * nrm[1]) + (q[2] - o[2]) * nrm[2];
  double a = fabs(h1), b = fabs(h2);
  double fp[3] = {p[0] - h1 * nrm[0], p[1] - h1 * nrm[1], p[2] - h1 * nrm[2]};
  double fq[3] = {q[0] - h2 * nrm[0], q[1] - h2 * nrm[1], q[2] - h2 * nrm[2]};
  double w[3] = {fq[0] - fp[0], fq[1] - fp[1], fq[2] - fp[2]};
  double L = sqrt64((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  double Lm = L > kTiny ? L : kTiny;
  double iLm = rcp64(Lm);
  double u[3] = {w[0] * iLm, w[1] * iLm, w[2] * iLm};
  double sm = (a + b) + L;
  double s = sm > kTiny ? sm : kTiny;
  double inv_s = rcp64(s);
  double ah = a * inv_s, bh = b * inv_s, Lh = L * inv_s;
  bool swap = n1 > n2;
  double ratio = (n1 < n2 ? n1 : n2) * rcp64(n1 > n2 ? n1 : n2);
  double ar = swap ? bh : ah;
  double bd = swap ? ah : bh;
  // float32 Newton in sigma (same safeguarded iteration)
  float arf = (float)ar, bdf = (float)bd, Lhf = (float)Lh, rf = (float)ratio;
  float ab = (float)(ah + bh);
  float hi0 = 1.0f - 8.0f * 1.1920929e-7f;
  float sig = Lhf * f_rsq(fmaxf(Lhf * Lhf + ab * ab, 1e-30f));
  sig = fminf(sig, hi0);
  float lo = 0.0f, hi = hi0;
  bool settled = false;
  for (int it = 0; it < kLocalIters32; ++it) {
    float v = rf * sig;
    float c1 = fmaxf((1.0f - sig) * (1.0f + sig), 1e-30f);
    float c2 = fmaxf((1.0f - v) * (1.0f + v), 1e-30f);
    float q1 = f_rsq(c1), q2 = f_rsq(c2);
    float F = arf * sig * q1 + bdf * v * q2 - Lhf;
    bool le = F <= 0.0f;
    lo = le ? sig : lo;
    hi = le ? hi : sig;
    float Fp = arf * (q1 * q1 * q1) + (bdf * rf) * (q2 * q2 * q2);
    float sn = sig - F * f_rcp(fmaxf(Fp, 1e-30f));
    bool ok = (sn >= lo) && (sn <= hi) && (sn == sn);
    const float nx = ok ? sn : 0.5f * (lo + hi);
    const bool done = ok && fabsf(nx - sig) <= kF32Settle * sig;
    sig = nx;
    if (done) {  // settled to a few ulp by a Newton step
      settled = true;
      break;
    }
  }
  if (settled && sig < kSigFast) {
    // Crossing straight from the float32 root, no float64 Newton polish.
    // The optical path n1 s1 + n2 s2 is stationary in the crossing point on
    // the plane (Fermat), so an in-plane error d of the crossing changes it
    // by at most d^2 (n1 / a + n2 / b) / 2 (a, b: the end points' heights
    // over the plane, a <= s1, b <= s2).  At the settled root the residual
    // F is float32 rounding noise of its terms (each <= Lh) and of the
    // float32 inputs, well under 1e-6 Lh; since dt / dsigma = ar (1 -
    // sigma^2)^(-3/2) <= F', the crossing parameter t is off by at most
    // |F|, and t's own rounding adds a few ulp of Lh: d <= kPosRelErr L.
    // The legs' directions move by at most d / a and d / b (first order:
    // incidence and refraction cosines, departure direction), held to
    // kDirRelErr.  Where both bounds hold -- the path error within opl_tol,
    // the direction error within kDirRelErr -- the crossing is
    // placed on the plane in float64 (fp + s t u), s1 is formed in float64
    // for it, and the residual is taken as converged: the float32 iteration
    // settled by a Newton step on an interior root from the reference's
    // starting point, where the float64 iteration converges quadratically
    // far inside the reference's 25-step budget.  Elsewhere (near grazing,
    // short legs of high curvature, bisection endings) the float64 polish
    // below runs.
    const float c1f = fmaxf((1.0f - sig) * (1.0f + sig), 1e-30f);
    const float trf = arf * sig * f_rsq(c1f);
    const float ttf = swap ? Lhf - trf : trf;
    const float tf = fminf(fmaxf(ttf, 0.0f), Lhf);
    const float af = (float)a, bf = (float)b;
    const float dpos = kPosRelErr * (float)L;
    const float curv = ((float)n1 * bf + (float)n2 * af) * f_rcp(fmaxf(af * bf, 1e-30f));
    if (0.5f * curv * dpos * dpos <= opl_tol && dpos <= kDirRelErr * fminf(af, bf)) {
      const double t = (double)tf;
      const double kTiny2 = kTiny * kTiny;
      const double q1s = ah * ah + t * t;
      const double ir1 = rsq64(q1s > kTiny2 ? q1s : kTiny2);
      const double r1 = q1s > kTiny2 ? q1s * ir1 : kTiny;
      out.res = 0.0;
      out.s1 = s * r1;
      const double st = s * t;
      out.x[0] = fp[0] + st * u[0];
      out.x[1] = fp[1] + st * u[1];
      out.x[2] = fp[2] + st * u[2];
      out.c_in = (float)(ah * ir1);
      const float bhf = (float)bh, d2f = Lhf - tf;
      out.c_out = bhf * f_rsq(fmaxf(bhf * bhf + d2f * d2f, 1e-30f));
      out.valid = h1 * h2 < 0.0;
      out.ambiguous = false;
      return;
    }
  }
  // float64 polish (plain Newton, clamped into [0, hi0)): one step, and
  // more (at most kPolishMax) while the Snell residual sits in the
  // ambiguous band.  The residual's sensitivity to sigma grows as the short
  // leg shrinks (1/bh: a target a few metres below a 20 km air leg), where
  // one step from the float32 root leaves it at 1e-11..1e-10.
  double sg = (double)sig;
  const double hi064 = 1.0 - 8.0 * kEps64;
  const double kTiny2 = kTiny * kTiny;
  double t, d2, ir1, ir2, r1, res;
  for (int it = 0;; ++it) {
    {
      double v = ratio * sg;
      double c1 = (1.0 - sg) * (1.0 + sg);
      c1 = c1 > kTiny ? c1 : kTiny;
      double c2 = (1.0 - v) * (1.0 + v);
      c2 = c2 > kTiny ? c2 : kTiny;
      double q1 = rsq64(c1), q2 = rsq64(c2);
      double F = (ar * sg * q1 + bd * v * q2) - Lh;
      double Fp = ar * (q1 * q1 * q1) + (bd * ratio) * (q2 * q2 * q2);
      double sn = sg - F * rcp64(Fp > kTiny ? Fp : kTiny);
      sg = sn < 0.0 ? 0.0 : (sn > hi064 ? hi064 : sn);
    }
    double c1 = (1.0 - sg) * (1.0 + sg);
    c1 = c1 > kTiny ? c1 : kTiny;
    double tr = ar * sg * rsq64(c1);
    double tt = swap ? Lh - tr : tr;
    t = tt < 0.0 ? 0.0 : (tt > Lh ? Lh : tt);
    d2 = Lh - t;
    // r = sqrt(q) and 1/r from one refined reciprocal square root
    const double q1s = ah * ah + t * t, q2s = bh * bh + d2 * d2;
    ir1 = rsq64(q1s > kTiny2 ? q1s : kTiny2);
    ir2 = rsq64(q2s > kTiny2 ? q2s : kTiny2);
    r1 = q1s > kTiny2 ? q1s * ir1 : kTiny;
    res = n1 * t * ir1 - n2 * d2 * ir2;
    const double ar_ = fabs(res);
    if (it + 1 >= kPolishMax || !(ar_ > kResLo && ar_ < kResHi)) break;
  }
  out.res = res;
  out.s1 = s * r1;
  double st = s * t;
  out.x[0] = fp[0] + st * u[0];
  out.x[1] = fp[1] + st * u[1];
  out.x[2] = fp[2] + st * u[2];
  out.c_in = (float)(ah * ir1);
  out.c_out = (float)(bh * ir2);
  double ares = fabs(out.res);
  out.valid = (h1 * h2 < 0.0) && (ares <= 1e-9);
  out.ambiguous = (h1 * h2 < 0.0) && (ares > kResLo) && (ares < kResHi);
}

// Mean-plane anchor in float32, plane-relative coordinates: returns the
// crossing point minus the plane point (float32).
SS_HD void snell_anchor32(const float p[3], const float q[3], const float nrm[3],
                          float n1, float n2, float x_rel[3]) {
  float h1 = p[0] * nrm[0] + p[1] * nrm[1] + p[2] * nrm[2];
  float h2 = q[0] * nrm[0] + q[1] * nrm[1] + q[2] * nrm[2];
  float a = fabsf(h1), b = fabsf(h2);
  float fp[3] = {p[0] - h1 * nrm[0], p[1] - h1 * nrm[1], p[2] - h1 * nrm[2]};
  float w[3] = {q[0] - h2 * nrm[0] - fp[0], q[1] - h2 * nrm[1] - fp[1],
                q[2] - h2 * nrm[2] - fp[2]};
  float L2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  float L = f_sqrt(L2);
  float iLm = f_rcp(fmaxf(L, 1e-30f));
  float s = fmaxf(a + b + L, 1e-30f);
  float is = f_rcp(s);
  float ah = a * is, bh = b * is, Lh = L * is;
  bool swap = n1 > n2;
  float ratio = fminf(n1, n2) * f_rcp(fmaxf(n1, n2));
  float ar = swap ? bh : ah;
  float bd = swap ? ah : bh;
  float hi0 = 1.0f - 8.0f * 1.1920929e-7f;
  float ab = ah + bh;
  float sig = fminf(Lh * f_rsq(fmaxf(Lh * Lh + ab * ab, 1e-30f)), hi0);
  float lo = 0.0f, hi = hi0;
  for (int it = 0; it < kMeanIters32; ++it) {
    float v = ratio * sig;
    float c1 = fmaxf((1.0f - sig) * (1.0f + sig), 1e-30f);
    float c2 = fmaxf((1.0f - v) * (1.0f + v), 1e-30f);
    float q1 = f_rsq(c1), q2 = f_rsq(c2);
    float F = ar * sig * q1 + bd * v * q2 - Lh;
    bool le = F <= 0.0f;
    lo = le ? sig : lo;
    hi = le ? hi : sig;
    float Fp = ar * (q1 * q1 * q1) + (bd * ratio) * (q2 * q2 * q2);
    float sn = sig - F * f_rcp(fmaxf(Fp, 1e-30f));
    bool ok = (sn >= lo) && (sn <= hi) && (sn == sn);
    const float nx = ok ? sn : 0.5f * (lo + hi);
    const bool done = ok && fabsf(nx - sig) <= kF32Settle * sig;
    sig = nx;
    if (done) break;  // settled to a few ulp (the cell guard band is 2e-4 cells)
  }
  float c1 = fmaxf((1.0f - sig) * (1.0f + sig), 1e-30f);
  float tr = ar * sig * f_rsq(c1);
  float tt = swap ? Lh - tr : tr;
  float t = fminf(fmaxf(tt, 0.0f), Lh);
  float st = s * t;
  x_rel[0] = fp[0] + st * w[0] * iLm;
  x_rel[1] = fp[1] + st * w[1] * iLm;
  x_rel[2] = fp[2] + st * w[2] * iLm;
}

// ---------------------------------------------------------------------------
// Refracted pair
// ---------------------------------------------------------------------------
struct RefrConst {
  double k0, gamma, t0, dt, c, taper_s, sigma, lc, n_exp;
  int n_samples, flags, n_crossed, n_terms, acf;
  PatternArgs pat;
  const double* cst;
  const double* desc;    // (n_crossed, SS_IDESC_COUNT)
  const double* tables;  // (rows, 6)
  const double* legs;    // (3, n_crossed + 1): n, eps, att
  const float* mtab = nullptr;  // staged D_Phi term constants (LDS) or null
  int joint, k_pad, n_newton, n_backtrack;  // joint solver (pads first)
  // derived
  double k_bin, t0_dt, cyc_per_m;
  float fast_opl_tol = 0.0f;  // optical-path error (m) allowed to a polish-free crossing
};

// Path sums handed from a crossing solver to the radiometry stage.
struct PathSums {
  double cur[3];       // last crossing (start of the final leg)
  double x_first[3];   // first crossing (air-leg departure)
  bool valid;
  double opl, loss_db, sum_par, sum_perp, tau2, c_first, c_last;
};

struct RefrFacet {
  double c[3], n[3], A, e1[3], e2[3], gam, ph_re, ph_im, l1, l2, dif, dph_re, dph_im;
};

// Mixed-tier facet record: float64 centre (path lengths need it), the rest
// float32 (radiometry only).
struct RefrFacet32 {
  double c[3];
  float n[3], A, e1[3], e2[3], gam, ph_re, ph_im, l1, l2, dif, dph_re, dph_im;
};

SS_HD void to_f32(const RefrFacet& a, RefrFacet32& b) {
  for (int j = 0; j < 3; ++j) {
    b.c[j] = a.c[j];
    b.n[j] = (float)a.n[j];
    b.e1[j] = (float)a.e1[j];
    b.e2[j] = (float)a.e2[j];
  }
  b.A = (float)a.A;
  b.gam = (float)a.gam;
  b.ph_re = (float)a.ph_re;
  b.ph_im = (float)a.ph_im;
  b.l1 = (float)a.l1;
  b.l2 = (float)a.l2;
  b.dif = (float)a.dif;
  b.dph_re = (float)a.dph_re;
  b.dph_im = (float)a.dph_im;
}

// FAST (the mixed tier): the normal's 1/|n| by the refined reciprocal
// square root instead of an IEEE division of an IEEE square root.
template <bool FAST = false>
SS_HD void lookup_cell(const RefrConst& rc, int i, int row, int col, double pt[3],
                       double nr[3]) {
#pragma clang fp contract(off)
  const double* d = rc.desc + i * SS_IDESC_COUNT;
  int nx = (int)d[SS_IDESC_NX];
  long off = (long)d[SS_IDESC_TABOFF] + (long)row * nx + col;
  const double* g = rc.tables + off * 6;
  pt[0] = g[0];
  pt[1] = g[1];
  pt[2] = g[2];
  double n0 = g[3], n1 = g[4], n2 = g[5];
  const double nn = (n0 * n0 + n1 * n1) + n2 * n2;
  double inn = FAST ? rsq64(nn) : 1.0 / sqrt(nn);
  nr[0] = n0 * inn;
  nr[1] = n1 * inn;
  nr[2] = n2 * inn;
}

SS_HD int clamp_round(double v, int hi) {
  double r = rint(v);  // round half to even, as np.round
  if (!(r > 0.0)) return 0;
  if (r > (double)hi) return hi;
  return (int)r;
}

// Float64 tier, sequential chain (kernels/multilayer.py:352-406).
SS_HD void seq_path_fp64(const RefrFacet& f, const TraceRec& t, const RefrConst& rc,
                         PathSums& ps) {
#pragma clang fp contract(off)
  const int flags = rc.flags;
  const double* nl = rc.legs;
  const double* epsl = rc.legs + (rc.n_crossed + 1);
  const double* attl = rc.legs + 2 * (rc.n_crossed + 1);
  double cur[3] = {t.p[0], t.p[1], t.p[2]};
  bool valid = true;
  double opl = 0.0, loss_db = 0.0, sum_par = 0.0, sum_perp = 0.0, tau2 = 1.0;
  double c_first = 1.0, c_last = 1.0, x_first[3] = {t.p[0], t.p[1], t.p[2]};
  for (int i = 0; i < rc.n_crossed; ++i) {
    const double* d = rc.desc + i * SS_IDESC_COUNT;
    double mp[3] = {d[SS_IDESC_MP], d[SS_IDESC_MP + 1], d[SS_IDESC_MP + 2]};
    double mn[3] = {d[SS_IDESC_MN], d[SS_IDESC_MN + 1], d[SS_IDESC_MN + 2]};
    Snell64 s;
    snell_fp64(cur, f.c, mp, mn, nl[i], nl[i + 1], kMeanIters, s);
    const double* cf = d + SS_IDESC_COEF;  // coef (3, 2) row-major
    double rc0 = (s.x[0] * cf[0] + s.x[1] * cf[2]) + cf[4];
    double rc1 = (s.x[0] * cf[1] + s.x[1] * cf[3]) + cf[5];
    int row = clamp_round(rc0, (int)d[SS_IDESC_NY] - 1);
    int col = clamp_round(rc1, (int)d[SS_IDESC_NX] - 1);
    double pt[3], nr[3];
    lookup_cell(rc, i, row, col, pt, nr);
    snell_fp64(cur, f.c, pt, nr, nl[i], nl[i + 1], kLocalIters, s);
    valid = valid && s.valid;
    double ci = cos(s.th1);
    double c_inc = ci > kCMin ? ci : kCMin;
    if (i == 0) {
      c_first = c_inc;
      x_first[0] = s.x[0];
      x_first[1] = s.x[1];
      x_first[2] = s.x[2];
    }
    {
      double e1 = epsl[i], e2 = epsl[i + 1];
      double n1 = sqrt(e1), n2 = sqrt(e2);
      double sin2 = (e1 / e2) * (1.0 - c_inc * c_inc);
      double q = 1.0 - sin2;
      double c2 = sqrt(q > 0.0 ? q : 0.0);
      double g = (n1 * c_inc - n2 * c2) / (n1 * c_inc + n2 * c2);
      tau2 = tau2 * (1.0 - g * g);
    }
    double co = cos(s.th2);
    if (flags & SS_CROSS_ROUGH) {
      double kt = rc.k0 * (nl[i] * c_inc - nl[i + 1] * co);
      double sk = d[SS_IDESC_SIGC] * kt;
      tau2 = tau2 * exp(-2.0 * (sk * sk));
    }
    opl = opl + nl[i] * s.s1;
    loss_db = loss_db + s.s1 * (attl[i] / 1000.0);
    sum_perp = sum_perp + s.s1 / nl[i];
    sum_par = sum_par + s.s1 / (nl[i] * c_inc * c_inc);
    c_last = co > kCMin ? co : kCMin;
    cur[0] = s.x[0];
    cur[1] = s.x[1];
    cur[2] = s.x[2];
  }
  for (int c = 0; c < 3; ++c) {
    ps.cur[c] = cur[c];
    ps.x_first[c] = x_first[c];
  }
  ps.valid = valid;
  ps.opl = opl;
  ps.loss_db = loss_db;
  ps.sum_par = sum_par;
  ps.sum_perp = sum_perp;
  ps.tau2 = tau2;
  ps.c_first = c_first;
  ps.c_last = c_last;
}

// Float64 tier, final leg + radiometry + bin (kernels/multilayer.py:500-583).
template <int RM = -1, int PAT = -1>
SS_HD void radiometry_fp64(const RefrFacet& f, const TraceRec& t, const RefrConst& rc,
                           const PathSums& ps, PairOut& o) {
#pragma clang fp contract(off)
  const int flags = rc.flags;
  const double* nl = rc.legs;
  const double* attl = rc.legs + 2 * (rc.n_crossed + 1);
  const double* cur = ps.cur;
  const double* x_first = ps.x_first;
  const bool valid = ps.valid;
  double opl = ps.opl, loss_db = ps.loss_db, sum_par = ps.sum_par, sum_perp = ps.sum_perp;
  const double tau2 = ps.tau2, c_first = ps.c_first, c_last = ps.c_last;
  const int J = rc.n_crossed;
  double dd[3] = {cur[0] - f.c[0], cur[1] - f.c[1], cur[2] - f.c[2]};
  double s_j = sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
  double inv = 1.0 / (s_j > 1e-30 ? s_j : 1e-30);
  double rh[3] = {dd[0] * inv, dd[1] * inv, dd[2] * inv};
  double nj = nl[J];
  opl = opl + nj * s_j;
  loss_db = loss_db + s_j * (attl[J] / 1000.0);
  sum_perp = sum_perp + s_j / nj;
  sum_par = sum_par + s_j / (nj * c_last * c_last);
  double l_perp = nl[0] * sum_perp;
  double l_par = nl[0] * c_first * c_first * sum_par;
  double flux = (nl[0] * c_first) / (nj * c_last);
  double att_f = pow(10.0, -loss_db / 10.0);
  double cos_t = (rh[0] * f.n[0] + rh[1] * f.n[1]) + rh[2] * f.n[2];
  double spread = tau2 * flux / (l_par * l_perp);
  if (PAT != 0 && rc.pat.kind != SS_PAT_ISOTROPIC) {
    double d0[3] = {x_first[0] - t.p[0], x_first[1] - t.p[1], x_first[2] - t.p[2]};
    double dm = sqrt((d0[0] * d0[0] + d0[1] * d0[1]) + d0[2] * d0[2]);
    dm = dm > 1e-30 ? dm : 1e-30;
    double dh[3] = {d0[0] / dm, d0[1] / dm, d0[2] / dm};
    double g = antenna_gain<double, PAT>(rc.pat, dh, t.pv0, t.pv1);
    spread = spread * (g * g);
  }
  double re, im, pwr;
  if (flags & SS_COHERENT) {
    double kj = rc.k0 * nj;
    double d1 = (rh[0] * f.e1[0] + rh[1] * f.e1[1]) + rh[2] * f.e1[2];
    double d2 = (rh[0] * f.e2[0] + rh[1] * f.e2[1]) + rh[2] * f.e2[2];
    double s1 = d_sinc(d1 * (kj / kPi));
    double s2 = d_sinc(d2 * (kj / kPi));
    double gam = (flags & SS_GAMMA_FACET) ? f.gam : rc.gamma;
    double amp = (kj / kTwoPi) * gam * cos_t * f.A * s1 * s2 * spread * att_f;
    if (flags & SS_TAPER) {
      double cc = cos_t < kCMin ? kCMin : (cos_t > 1.0 ? 1.0 : cos_t);
      double ct2 = cc * cc;
      amp = amp * exp(-(1.0 - ct2) / (ct2 * (2.0 * rc.taper_s * rc.taper_s)));
    }
    double are = amp, aim = 0.0;
    if (RM != 0 && (flags & SS_ROUGH)) {
      double kk = 2.0 * kj * cos_t;
      double sk = rc.sigma * kk;
      double matt = exp(-0.5 * (sk * sk));
      are = amp * matt;
      if (f.A > 0.0) {
        double dp = d_phi<double, RM>(rc.sigma, rc.lc, kk, 2.0 * kj * d1 / f.l1,
                                  2.0 * kj * d2 / f.l2, f.l1, f.l2, rc.n_terms,
                                  (flags & SS_D_PHI_AREA) != 0, rc.acf, rc.cst);
        double ai = (kj / kTwoPi) * gam * cos_t * spread * att_f * sqrt(dp);
        are = are + ai * f.ph_re;
        aim = aim + ai * f.ph_im;
      }
    }
    if ((flags & SS_DIFFUSE) && f.A > 0.0) {
      double ct = cos_t > 0.0 ? cos_t : 0.0;
      double a_d = sqrt(f.A / kTwoPi) * f.dif * ct * pow(ct, 0.5 * rc.n_exp) *
                   spread * att_f;
      are = are + a_d * f.dph_re;
      aim = aim + a_d * f.dph_im;
    }
    double phase = -2.0 * rc.k0 * opl;
    double sp = sin(phase), cp = cos(phase);
    re = -(are * sp + aim * cp);
    im = are * cp - aim * sp;
    pwr = re * re + im * im;
  } else {
    double a = f.A * cos_t * spread;
    re = a * a * att_f * att_f;
    im = 0.0;
    pwr = re;
    o.power = true;
  }
  double pos = (2.0 * opl / rc.c - rc.t0) / rc.dt;
  double fb = floor(pos);
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.nparts = 1;
  o.re[0] = re;
  o.im[0] = im;
  o.bin[0] = b;
  o.valid = valid;
  o.raw_bin = valid ? b : SS_BIN_INVALID;
  o.pwr_invalid = valid ? 0.0 : pwr;
  o.right = (((f.c[0] - t.p[0]) * t.u[0] + (f.c[1] - t.p[1]) * t.u[1]) +
             (f.c[2] - t.p[2]) * t.u[2]) > 0.0
                ? 1
                : 0;
}

// Mixed tier, sequential chain.  Returns false when the pair must be redone
// on the float64 tier (ambiguous Snell residual).
template <typename FT>
SS_HD bool seq_path_mixed(const FT& f, const TraceRec& t, const RefrConst& rc,
                          PathSums& ps) {
  const int flags = rc.flags;
  const double* nl = rc.legs;
  const double* epsl = rc.legs + (rc.n_crossed + 1);
  const double* attl = rc.legs + 2 * (rc.n_crossed + 1);
  double cur[3] = {t.p[0], t.p[1], t.p[2]};
  bool valid = true;
  double opl = 0.0;
  float loss_db = 0.0f, sum_par = 0.0f, sum_perp = 0.0f, tau2 = 1.0f;
  float c_first = 1.0f, c_last = 1.0f;
  double x_first[3] = {t.p[0], t.p[1], t.p[2]};
  for (int i = 0; i < rc.n_crossed; ++i) {
    const double* d = rc.desc + i * SS_IDESC_COUNT;
    const double* cf = d + SS_IDESC_COEF;
    // anchor: float32 in coordinates relative to the mean-plane point
    float pr[3] = {(float)(cur[0] - d[SS_IDESC_MP]), (float)(cur[1] - d[SS_IDESC_MP + 1]),
                   (float)(cur[2] - d[SS_IDESC_MP + 2])};
    float qr[3] = {(float)(f.c[0] - d[SS_IDESC_MP]), (float)(f.c[1] - d[SS_IDESC_MP + 1]),
                   (float)(f.c[2] - d[SS_IDESC_MP + 2])};
    float mn[3] = {(float)d[SS_IDESC_MN], (float)d[SS_IDESC_MN + 1], (float)d[SS_IDESC_MN + 2]};
    float xr[3];
    snell_anchor32(pr, qr, mn, (float)nl[i], (float)nl[i + 1], xr);
    double x0 = (double)xr[0] + d[SS_IDESC_MP], x1 = (double)xr[1] + d[SS_IDESC_MP + 1];
    double rc0 = (x0 * cf[0] + x1 * cf[2]) + cf[4];
    double rc1 = (x0 * cf[1] + x1 * cf[3]) + cf[5];
    double f0 = rc0 - floor(rc0), f1 = rc1 - floor(rc1);
    if (fabs(f0 - 0.5) < kCellGuard || fabs(f1 - 0.5) < kCellGuard) {
      // near a lookup rounding boundary: the reference float64 anchor
      double mp[3] = {d[SS_IDESC_MP], d[SS_IDESC_MP + 1], d[SS_IDESC_MP + 2]};
      double mnd[3] = {d[SS_IDESC_MN], d[SS_IDESC_MN + 1], d[SS_IDESC_MN + 2]};
      Snell64 s;
      snell_fp64(cur, f.c, mp, mnd, nl[i], nl[i + 1], kMeanIters, s);
      {
#pragma clang fp contract(off)
        rc0 = (s.x[0] * cf[0] + s.x[1] * cf[2]) + cf[4];
        rc1 = (s.x[0] * cf[1] + s.x[1] * cf[3]) + cf[5];
      }
    }
    int row = clamp_round(rc0, (int)d[SS_IDESC_NY] - 1);
    int col = clamp_round(rc1, (int)d[SS_IDESC_NX] - 1);
    double pt[3], nr[3];
    lookup_cell<true>(rc, i, row, col, pt, nr);
    SnellMixed s;
    snell_mixed(cur, f.c, pt, nr, nl[i], nl[i + 1], rc.fast_opl_tol, s);
    if (s.ambiguous) return false;
    valid = valid && s.valid;
    float c_inc = fmaxf(s.c_in, (float)kCMin);
    if (i == 0) {
      c_first = c_inc;
      x_first[0] = s.x[0];
      x_first[1] = s.x[1];
      x_first[2] = s.x[2];
    }
    {
      float e1 = (float)epsl[i], e2 = (float)epsl[i + 1];
      float n1 = (float)nl[i], n2 = (float)nl[i + 1];
      float sin2 = (e1 * f_rcp(e2)) * (1.0f - c_inc * c_inc);
      float c2 = f_sqrt(fmaxf(1.0f - sin2, 0.0f));
      float g = (n1 * c_inc - n2 * c2) * f_rcp(n1 * c_inc + n2 * c2);
      tau2 *= (1.0f - g * g);
    }
    float co = s.c_out;
    if (flags & SS_CROSS_ROUGH) {
      float kt = (float)rc.k0 * ((float)nl[i] * c_inc - (float)nl[i + 1] * co);
      float sk = (float)d[SS_IDESC_SIGC] * kt;
      tau2 *= f_exp(-2.0f * sk * sk);
    }
    float s1f = (float)s.s1;
    float ni = (float)nl[i];
    opl = opl + nl[i] * s.s1;
    loss_db += s1f * ((float)attl[i] * 1e-3f);
    const float ini = f_rcp(ni);
    sum_perp += s1f * ini;
    sum_par += s1f * ini * f_rcp(c_inc * c_inc);
    c_last = fmaxf(co, (float)kCMin);
    cur[0] = s.x[0];
    cur[1] = s.x[1];
    cur[2] = s.x[2];
  }
  for (int c = 0; c < 3; ++c) {
    ps.cur[c] = cur[c];
    ps.x_first[c] = x_first[c];
  }
  ps.valid = valid;
  ps.opl = opl;
  ps.loss_db = loss_db;
  ps.sum_par = sum_par;
  ps.sum_perp = sum_perp;
  ps.tau2 = tau2;
  ps.c_first = c_first;
  ps.c_last = c_last;
  return true;
}

// Mixed tier, final leg + radiometry + bin.  Returns false when the bin
// position lies within kBinGuard of an edge (the caller redoes the pair on
// the float64 tier).
template <int RM = -1, int PAT = -1, typename FT = RefrFacet, bool STAGED = false>
SS_HD bool radiometry_mixed(const FT& f, const TraceRec& t, const RefrConst& rc,
                            const PathSums& ps, PairOut& o) {
  const int flags = rc.flags;
  const double* nl = rc.legs;
  const double* attl = rc.legs + 2 * (rc.n_crossed + 1);
  const double* cur = ps.cur;
  const double* x_first = ps.x_first;
  const bool valid = ps.valid;
  double opl = ps.opl;
  float loss_db = (float)ps.loss_db, sum_par = (float)ps.sum_par,
        sum_perp = (float)ps.sum_perp, tau2 = (float)ps.tau2;
  const float c_first = (float)ps.c_first, c_last = (float)ps.c_last;
  const int J = rc.n_crossed;
  double dd[3] = {cur[0] - f.c[0], cur[1] - f.c[1], cur[2] - f.c[2]};
  double s_j = sqrt64(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
  double nj = nl[J];
  opl = opl + nj * s_j;
  // bin position (fast) -- redo the pair on the float64 tier near an edge
  double pos = fma(opl, rc.k_bin, -rc.t0_dt);
  double fb = floor(pos);
  double fr = pos - fb;
  if (fr < kBinGuard || fr > 1.0 - kBinGuard) return false;
  float sjf = (float)s_j;
  float isj = f_rcp(fmaxf(sjf, 1e-30f));
  float rh[3] = {(float)dd[0] * isj, (float)dd[1] * isj, (float)dd[2] * isj};
  float njf = (float)nj;
  loss_db += sjf * ((float)attl[J] * 1e-3f);
  const float injf = f_rcp(njf);
  sum_perp += sjf * injf;
  sum_par += sjf * injf * f_rcp(c_last * c_last);
  float n0 = (float)nl[0];
  float l_perp = n0 * sum_perp;
  float l_par = n0 * c_first * c_first * sum_par;
  float flux = (n0 * c_first) * injf * f_rcp(c_last);
  float att_f = f_exp2(-loss_db * (3.321928094887362f / 10.0f));  // 10^(-loss/10)
  float cos_t = rh[0] * (float)f.n[0] + rh[1] * (float)f.n[1] + rh[2] * (float)f.n[2];
  float spread = tau2 * flux * f_rcp(l_par * l_perp);
  if (PAT != 0 && rc.pat.kind != SS_PAT_ISOTROPIC) {
    float d0[3] = {(float)(x_first[0] - t.p[0]), (float)(x_first[1] - t.p[1]),
                   (float)(x_first[2] - t.p[2])};
    float dm = fmaxf(f_sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]), 1e-30f);
    float idm = f_rcp(dm);
    float dh[3] = {d0[0] * idm, d0[1] * idm, d0[2] * idm};
    float pv0[3] = {(float)t.pv0[0], (float)t.pv0[1], (float)t.pv0[2]};
    float pv1[3] = {(float)t.pv1[0], (float)t.pv1[1], (float)t.pv1[2]};
    float g = antenna_gain<float, PAT, STAGED>(rc.pat, dh, pv0, pv1);
    spread *= g * g;
  }
  float re, im, pwr;
  if (flags & SS_COHERENT) {
    float kj = (float)(rc.k0 * nj);
    float d1 = rh[0] * (float)f.e1[0] + rh[1] * (float)f.e1[1] + rh[2] * (float)f.e1[2];
    float d2 = rh[0] * (float)f.e2[0] + rh[1] * (float)f.e2[1] + rh[2] * (float)f.e2[2];
    float kop = kj * (float)(1.0 / kPi);
    float s1 = f_sinc(d1 * kop);
    float s2 = f_sinc(d2 * kop);
    float gam = (flags & SS_GAMMA_FACET) ? (float)f.gam : (float)rc.gamma;
    float base = kj * (float)(1.0 / kTwoPi) * gam * cos_t * spread * att_f;
    float amp = base * (float)f.A * s1 * s2;
    if (flags & SS_TAPER) {
      float cc = fminf(fmaxf(cos_t, (float)kCMin), 1.0f);
      float ct2 = cc * cc;
      float ts = (float)rc.taper_s;
      amp *= f_exp(-(1.0f - ct2) * f_rcp(ct2 * (2.0f * ts * ts)));
    }
    float are = amp, aim = 0.0f;
    if (RM != 0 && (flags & SS_ROUGH)) {
      float kk = 2.0f * kj * cos_t;
      float sk = (float)rc.sigma * kk;
      are = amp * f_exp(-0.5f * sk * sk);
      if (f.A > 0.0) {
        float l1 = (float)f.l1, l2 = (float)f.l2;
        float dp = d_phi<float, RM, true>((float)rc.sigma, (float)rc.lc, kk, 2.0f * kj * d1 * f_rcp(l1),
                                2.0f * kj * d2 * f_rcp(l2), l1, l2, rc.n_terms,
                                (flags & SS_D_PHI_AREA) != 0, rc.acf, rc.cst, rc.mtab);
        float ai = base * f_sqrt(dp);
        are += ai * (float)f.ph_re;
        aim += ai * (float)f.ph_im;
      }
    }
    if ((flags & SS_DIFFUSE) && f.A > 0.0) {
      float ct = fmaxf(cos_t, 0.0f);
      float a_d = f_sqrt((float)f.A * (float)(1.0 / kTwoPi)) * (float)f.dif * ct *
                  t_pow<float>(ct, 0.5f * (float)rc.n_exp) * spread * att_f;
      are += a_d * (float)f.dph_re;
      aim += a_d * (float)f.dph_im;
    }
    double cyc = opl * rc.cyc_per_m;
    float ph = (float)(cyc - rint(cyc));
    float sp = -f_sin_rev(ph), cp = f_cos_rev(ph);
    re = -(are * sp + aim * cp);
    im = are * cp - aim * sp;
    pwr = re * re + im * im;
  } else {
    float a = (float)f.A * cos_t * spread;
    re = a * a * att_f * att_f;
    im = 0.0f;
    pwr = re;
    o.power = true;
  }
  int b = (fb < -2147483647.0) ? (-2147483647) : (fb > 2147483646.0 ? 2147483646 : (int)fb);
  o.nparts = 1;
  o.re[0] = re;
  o.im[0] = im;
  o.bin[0] = b;
  o.valid = valid;
  o.raw_bin = valid ? b : SS_BIN_INVALID;
  o.pwr_invalid = valid ? 0.0 : (double)pwr;
  o.right = (((f.c[0] - t.p[0]) * t.u[0] + (f.c[1] - t.p[1]) * t.u[1]) +
             (f.c[2] - t.p[2]) * t.u[2]) > 0.0
                ? 1
                : 0;
  return true;
}

}  // namespace ss

/*
 * oracle.c -- float64 C restatement of the reference's surface and
 * sequential-refracted coherent kernels, TEST INFRASTRUCTURE ONLY: it is
 * the timed CPU baseline (OpenMP over traces) and is cross-checked against
 * the NumPy oracle (oracle/coherent.py, oracle/multilayer.py).
 *
 * Follows soundersim kernels/coherent.py:107-128 (LPA field), :94-104
 * (taper), roughness.py:83-186 (area-only D_Phi, Gaussian and exponential
 * ACF), antenna.py:204-258 (dipole, tabulated), geometry.py:18-24 (bins),
 * kernels/multilayer.py:196-266 (two-point Snell), :352-406 (sequential
 * chain), :500-583 (radiometry).  Feature subset: the full finite-facet
 * Gaussian D_Phi, array patterns, split sides and interp bins are served
 * by the NumPy oracle only.
 *
 * bins (nullable): (T, F) int32, the floor bin of every pair clamped to
 * [-2147483647, 2147483646] as the kernels report it, or INT32_MIN for a
 * refracted pair whose path is invalid.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846
#define TWO_PI 6.28318530717958647692
#define TINY 1e-30
#define CMIN 1e-9

static int32_t clamp_bin(double fb) {
  return fb < -2147483647.0 ? -2147483647 : (fb > 2147483646.0 ? 2147483646 : (int32_t)fb);
}

static double sincn(double x) {
  double y = PI * (x == 0.0 ? 1.0e-20 : x);
  return sin(y) / y;
}

static double interp(double x, const double* xp, const double* fp, int n) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  return (fp[hi] - fp[lo]) / (xp[hi] - xp[lo]) * (x - xp[lo]) + fp[lo];
}

/* pattern: 0 isotropic, 1 dipole, 5 tabulated (tab = theta[n] | gain[n]) */
static double gain(int kind, const double* dh, const double* pv, const double* tab, int ntab) {
  double ca = (dh[0] * pv[0] + dh[1] * pv[1]) + dh[2] * pv[2];
  if (kind == 1) {
    double s = sqrt(fmax(1.0 - ca * ca, 1e-12));
    return cos((PI / 2.0) * ca) / s;
  }
  if (kind == 5) {
    double ct = ca < -1.0 ? -1.0 : (ca > 1.0 ? 1.0 : ca);
    return interp(acos(ct), tab, tab + ntab, ntab);
  }
  return 1.0;
}

/* area-only D_Phi; acf 0 Gaussian, 1 exponential */
static double dphi_area(double sigma, double l, double K, double A0, double B0, double Lx,
                        double Ly, int n_terms, int acf) {
  double x = (sigma * K) * (sigma * K);
  if (!(x > 0.0)) return 0.0;
  double logx = log(x), acc = 0.0;
  for (int m = 1; m <= n_terms; ++m) {
    double mf = (double)m;
    double logp = mf * logx - lgamma(mf + 1.0) - x;
    double term;
    if (acf == 1) {
      double kb2 = A0 * A0 + B0 * B0, q = (l / mf) * (l / mf);
      term = exp(logp) * (2.0 * PI * q * pow(1.0 + kb2 * q, -1.5)) * Lx * Ly;
    } else {
      term = exp(logp - (A0 * A0 + B0 * B0) * (l * l) / (4.0 * mf)) * (PI * (l * l) / mf) * Lx * Ly;
    }
    acc += term;
  }
  return acc > 0.0 ? acc : 0.0;
}

typedef struct {
  double k, gamma, t0, dt, c, taper_s, sigma, lc;
  long n;
  int n_terms, acf, pat, ntab, nthreads;
  const double* tab;
} oc_params;

int oracle_surface(const oc_params* P, const double* pos, const double* pv, const double* rref,
                   long T, const double* cen, const double* nrm, const double* area,
                   const double* e1, const double* e2, const double* gf, const double* ph,
                   long F, double* field, double* dropped, int32_t* bins) {
  memset(field, 0, sizeof(double) * T * P->n * 2);
  memset(dropped, 0, sizeof(double) * T);
#pragma omp parallel for schedule(dynamic, 1) num_threads(P->nthreads)
  for (long t = 0; t < T; ++t) {
    const double* p = pos + 3 * t;
    double* h = field + 2 * P->n * t;
    double drop = 0.0;
    for (long f = 0; f < F; ++f) {
      const double* cf = cen + 3 * f;
      double d[3] = {p[0] - cf[0], p[1] - cf[1], p[2] - cf[2]};
      double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
      const double* nf = nrm + 3 * f;
      double cs = ((d[0] * nf[0] + d[1] * nf[1]) + d[2] * nf[2]) / r;
      double rh[3] = {d[0] / r, d[1] / r, d[2] / r};
      const double *a1 = e1 + 3 * f, *a2 = e2 + 3 * f;
      double d1 = (rh[0] * a1[0] + rh[1] * a1[1]) + rh[2] * a1[2];
      double d2 = (rh[0] * a2[0] + rh[1] * a2[1]) + rh[2] * a2[2];
      double s1 = sincn(d1 * (P->k / PI)), s2 = sincn(d2 * (P->k / PI));
      double g = gf ? gf[f] : P->gamma;
      double amp = (P->k / TWO_PI) * g * cs * area[f] * s1 * s2 / (r * r);
      if (P->taper_s > 0.0) {
        double cc = cs < 1e-9 ? 1e-9 : (cs > 1.0 ? 1.0 : cs), c2 = cc * cc;
        amp = amp * exp(-(1.0 - c2) / (c2 * (2.0 * P->taper_s * P->taper_s)));
      }
      double phase = (-2.0 * P->k) * (r - rref[t]);
      double cre = amp, cim = 0.0;
      if (P->acf >= 0 && ph) {
        double l1 = sqrt((a1[0] * a1[0] + a1[1] * a1[1]) + a1[2] * a1[2]);
        double l2 = sqrt((a2[0] * a2[0] + a2[1] * a2[1]) + a2[2] * a2[2]);
        double kk = 2.0 * P->k * cs;
        double dp = area[f] > 0 ? dphi_area(P->sigma, P->lc, kk, 2.0 * P->k * d1 / l1,
                                            2.0 * P->k * d2 / l2, l1, l2, P->n_terms, P->acf)
                                : 0.0;
        double ai = area[f] > 0 ? (P->k / TWO_PI) * g * cs / (r * r) * sqrt(dp) : 0.0;
        double sk = P->sigma * kk;
        cre = amp * exp(-0.5 * (sk * sk)) + ai * ph[2 * f];
        cim = ai * ph[2 * f + 1];
      }
      double sp = sin(phase), cp = cos(phase);
      double re = -(cre * sp + cim * cp), im = cre * cp - cim * sp;
      if (P->pat) {
        double dh[3] = {-rh[0], -rh[1], -rh[2]};
        double gg = gain(P->pat, dh, pv + 3 * t, P->tab, P->ntab);
        re *= gg * gg;
        im *= gg * gg;
      }
      double b = floor((2.0 * r / P->c - P->t0) / P->dt);
      if (bins) bins[t * F + f] = clamp_bin(b);
      if (b >= 0.0 && b < (double)P->n) {
        h[2 * (long)b] += re;
        h[2 * (long)b + 1] += im;
      } else {
        drop += re * re + im * im;
      }
    }
    dropped[t] = drop;
    double th = fmod(2.0 * P->k * rref[t], TWO_PI), c = cos(-th), s = sin(-th);
    for (long i = 0; i < P->n; ++i) {
      double a = h[2 * i], bb = h[2 * i + 1];
      h[2 * i] = a * c - bb * s;
      h[2 * i + 1] = a * s + bb * c;
    }
  }
  return 0;
}

typedef struct {
  double x[3], th1, th2, s1;
  int valid;
} snell_t;

static void snell(const double* p, const double* q, const double* o, const double* nr, double n1,
                  double n2, int n_iter, snell_t* out) {
  const double eps = 2.220446049250313e-16;
  double h1 = ((p[0] - o[0]) * nr[0] + (p[1] - o[1]) * nr[1]) + (p[2] - o[2]) * nr[2];
  double h2 = ((q[0] - o[0]) * nr[0] + (q[1] - o[1]) * nr[1]) + (q[2] - o[2]) * nr[2];
  double a = fabs(h1), b = fabs(h2);
  double fp[3], fq[3], w[3], u[3];
  for (int i = 0; i < 3; ++i) {
    fp[i] = p[i] - h1 * nr[i];
    fq[i] = q[i] - h2 * nr[i];
    w[i] = fq[i] - fp[i];
  }
  double L = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  double Lm = L > TINY ? L : TINY;
  for (int i = 0; i < 3; ++i) u[i] = w[i] / Lm;
  double s = (a + b) + L;
  s = s > TINY ? s : TINY;
  double is = 1.0 / s, ah = a * is, bh = b * is, Lh = L * is;
  int swap = n1 > n2;
  double ratio = fmin(n1, n2) / fmax(n1, n2);
  double ar = swap ? bh : ah, bd = swap ? ah : bh;
  double hi0 = 1.0 - 8.0 * eps, ab = ah + bh;
  double den = sqrt(Lh * Lh + ab * ab);
  double sig = fmin(Lh / (den > TINY ? den : TINY), hi0), lo = 0.0, hi = hi0;
  for (int it = 0; it < n_iter; ++it) {
    double v = ratio * sig;
    double c1 = fmax((1.0 - sig) * (1.0 + sig), TINY), c2 = fmax((1.0 - v) * (1.0 + v), TINY);
    double F = (ar * sig * (1.0 / sqrt(c1)) + bd * v * (1.0 / sqrt(c2))) - Lh;
    if (F <= 0.0) lo = sig; else hi = sig;
    double Fp = ar / (c1 * sqrt(c1)) + (bd * ratio) / (c2 * sqrt(c2));
    double sn = sig - F / (Fp > TINY ? Fp : TINY);
    sig = (sn >= lo && sn <= hi && isfinite(sn)) ? sn : 0.5 * (lo + hi);
  }
  double c1 = fmax((1.0 - sig) * (1.0 + sig), TINY);
  double tr = ar * sig * (1.0 / sqrt(c1));
  double tt = swap ? Lh - tr : tr;
  double t = tt < 0.0 ? 0.0 : (tt > Lh ? Lh : tt), d2 = Lh - t;
  double r1 = fmax(sqrt(ah * ah + t * t), TINY), r2 = fmax(sqrt(bh * bh + d2 * d2), TINY);
  double res = n1 * t / r1 - n2 * d2 / r2;
  out->s1 = s * r1;
  for (int i = 0; i < 3; ++i) out->x[i] = fp[i] + (s * t) * u[i];
  out->th1 = atan2(t, ah);
  out->th2 = atan2(d2, bh);
  out->valid = (h1 * h2 < 0.0) && (fabs(res) <= 1e-9);
}

/* crossed interface i: tab (ny*nx, 6), desc (16): ny nx coef[6] mp[3] mn[3] sigc taboff */
int oracle_refracted(const oc_params* P, const double* pos, const double* pv, long T,
                     const double* cen, const double* nrm, const double* area, const double* e1,
                     const double* e2, const double* gf, const double* ph, long F,
                     int n_crossed, const double* desc, const double* tables,
                     const double* n_leg, const double* eps_leg, const double* att_leg,
                     double* field, double* dropped, int32_t* bins) {
  memset(field, 0, sizeof(double) * T * P->n * 2);
  memset(dropped, 0, sizeof(double) * T);
#pragma omp parallel for schedule(dynamic, 1) num_threads(P->nthreads)
  for (long t = 0; t < T; ++t) {
    const double* p = pos + 3 * t;
    double* h = field + 2 * P->n * t;
    double drop = 0.0;
    for (long f = 0; f < F; ++f) {
      const double* q = cen + 3 * f;
      double cur[3] = {p[0], p[1], p[2]}, xf[3] = {p[0], p[1], p[2]};
      int valid = 1;
      double opl = 0.0, loss = 0.0, spar = 0.0, sperp = 0.0, tau2 = 1.0, c0 = 1.0, cl = 1.0;
      for (int i = 0; i < n_crossed; ++i) {
        const double* d = desc + 16 * i;
        snell_t s;
        snell(cur, q, d + 8, d + 11, n_leg[i], n_leg[i + 1], 10, &s);
        double rc0 = (s.x[0] * d[2] + s.x[1] * d[4]) + d[6];
        double rc1 = (s.x[0] * d[3] + s.x[1] * d[5]) + d[7];
        int ny = (int)d[0], nx = (int)d[1];
        double rr = rint(rc0), cc = rint(rc1);
        int row = rr > 0 ? (rr > ny - 1 ? ny - 1 : (int)rr) : 0;
        int col = cc > 0 ? (cc > nx - 1 ? nx - 1 : (int)cc) : 0;
        const double* g6 = tables + 6 * ((long)d[15] + (long)row * nx + col);
        double inn = 1.0 / sqrt((g6[3] * g6[3] + g6[4] * g6[4]) + g6[5] * g6[5]);
        double nr[3] = {g6[3] * inn, g6[4] * inn, g6[5] * inn};
        snell(cur, q, g6, nr, n_leg[i], n_leg[i + 1], 25, &s);
        valid = valid && s.valid;
        double ci = fmax(cos(s.th1), CMIN);
        if (i == 0) {
          c0 = ci;
          memcpy(xf, s.x, sizeof xf);
        }
        double n1 = sqrt(eps_leg[i]), n2 = sqrt(eps_leg[i + 1]);
        double sin2 = (eps_leg[i] / eps_leg[i + 1]) * (1.0 - ci * ci);
        double c2 = sqrt(fmax(1.0 - sin2, 0.0));
        double gg = (n1 * ci - n2 * c2) / (n1 * ci + n2 * c2);
        tau2 = tau2 * (1.0 - gg * gg);
        opl = opl + n_leg[i] * s.s1;
        loss = loss + s.s1 * (att_leg[i] / 1000.0);
        sperp = sperp + s.s1 / n_leg[i];
        spar = spar + s.s1 / (n_leg[i] * ci * ci);
        cl = fmax(cos(s.th2), CMIN);
        memcpy(cur, s.x, sizeof cur);
      }
      double dd[3] = {cur[0] - q[0], cur[1] - q[1], cur[2] - q[2]};
      double sj = sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
      double inv = 1.0 / fmax(sj, 1e-30);
      double rh[3] = {dd[0] * inv, dd[1] * inv, dd[2] * inv};
      double nj = n_leg[n_crossed];
      opl = opl + nj * sj;
      loss = loss + sj * (att_leg[n_crossed] / 1000.0);
      sperp = sperp + sj / nj;
      spar = spar + sj / (nj * cl * cl);
      double lperp = n_leg[0] * sperp, lpar = n_leg[0] * c0 * c0 * spar;
      double flux = (n_leg[0] * c0) / (nj * cl);
      double attf = pow(10.0, -loss / 10.0);
      const double* nf = nrm + 3 * f;
      double cost = (rh[0] * nf[0] + rh[1] * nf[1]) + rh[2] * nf[2];
      double spread = tau2 * flux / (lpar * lperp);
      if (P->pat) {
        double d0[3] = {xf[0] - p[0], xf[1] - p[1], xf[2] - p[2]};
        double dm = fmax(sqrt((d0[0] * d0[0] + d0[1] * d0[1]) + d0[2] * d0[2]), 1e-30);
        double dh[3] = {d0[0] / dm, d0[1] / dm, d0[2] / dm};
        double gg = gain(P->pat, dh, pv + 3 * t, P->tab, P->ntab);
        spread = spread * (gg * gg);
      }
      double kj = P->k * nj;
      const double *a1 = e1 + 3 * f, *a2 = e2 + 3 * f;
      double d1 = (rh[0] * a1[0] + rh[1] * a1[1]) + rh[2] * a1[2];
      double d2 = (rh[0] * a2[0] + rh[1] * a2[1]) + rh[2] * a2[2];
      double g = gf ? gf[f] : P->gamma;
      double amp = (kj / TWO_PI) * g * cost * area[f] * sincn(d1 * (kj / PI)) *
                   sincn(d2 * (kj / PI)) * spread * attf;
      if (P->taper_s > 0.0) {
        double cc = cost < CMIN ? CMIN : (cost > 1.0 ? 1.0 : cost), ct2 = cc * cc;
        amp = amp * exp(-(1.0 - ct2) / (ct2 * (2.0 * P->taper_s * P->taper_s)));
      }
      double are = amp, aim = 0.0;
      if (P->acf >= 0 && ph) {
        double kk = 2.0 * kj * cost, sk = P->sigma * kk;
        are = amp * exp(-0.5 * (sk * sk));
        if (area[f] > 0) {
          double l1 = sqrt((a1[0] * a1[0] + a1[1] * a1[1]) + a1[2] * a1[2]);
          double l2 = sqrt((a2[0] * a2[0] + a2[1] * a2[1]) + a2[2] * a2[2]);
          double dp = dphi_area(P->sigma, P->lc, kk, 2.0 * kj * d1 / l1, 2.0 * kj * d2 / l2,
                                l1, l2, P->n_terms, P->acf);
          double ai = (kj / TWO_PI) * g * cost * spread * attf * sqrt(dp);
          are += ai * ph[2 * f];
          aim += ai * ph[2 * f + 1];
        }
      }
      double phase = -2.0 * P->k * opl, sp = sin(phase), cp = cos(phase);
      double re = -(are * sp + aim * cp), im = are * cp - aim * sp;
      double b = floor((2.0 * opl / P->c - P->t0) / P->dt);
      if (bins) bins[t * F + f] = valid ? clamp_bin(b) : INT32_MIN;
      if (valid && b >= 0.0 && b < (double)P->n) {
        h[2 * (long)b] += re;
        h[2 * (long)b + 1] += im;
      } else {
        drop += re * re + im * im;
      }
    }
    dropped[t] = drop;
  }
  return 0;
}

// ss_joint.h -- joint N-interface refraction solve per (trace, facet) pair,
// float64 (refraction_joint.py:109-322, kernels/multilayer.py:408-457).
//
// Unknowns: two in-plane offsets per crossed local plane; residual
// F_i = B_i^T (n_i d_i - n_{i+1} d_{i+1}); Jacobian block-tridiagonal with
// A_i = B_i^T (W_i + W_{i+1}) B_i + mu I, C_i = -B_i^T W_{i+1} B_{i+1},
// W_k = n_k (I - d_k d_k^T) / s_k, solved by block Thomas elimination; step
// halving on |F|^2; initialised by the sequential two-point chain; scaled by
// |p - q|.  The stack (real interfaces plus index-matched air pads on top) is
// at most NMAX deep; per-lane state lives in private arrays.
#pragma once

#include "ss_refract.h"

namespace ss {

struct JointState {
  static constexpr int kMax = 16;
};

SS_HD void mat3_sym_outer(const double d[3], double w, double out[6]) {
  // w (I - d d^T) stored as xx, xy, xz, yy, yz, zz
  out[0] = w * (1.0 - d[0] * d[0]);
  out[1] = w * (0.0 - d[0] * d[1]);
  out[2] = w * (0.0 - d[0] * d[2]);
  out[3] = w * (1.0 - d[1] * d[1]);
  out[4] = w * (0.0 - d[1] * d[2]);
  out[5] = w * (1.0 - d[2] * d[2]);
}

SS_HD void sym_mul_vec(const double m[6], const double v[3], double o[3]) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[1] * v[0] + m[3] * v[1] + m[4] * v[2];
  o[2] = m[2] * v[0] + m[4] * v[1] + m[5] * v[2];
}

// r = Bi^T M Bj (2x2); B stored column-major as b[col][xyz]
SS_HD void btmb(const double bi[2][3], const double m[6], const double bj[2][3],
                double r[2][2]) {
  for (int c = 0; c < 2; ++c) {
    double mv[3];
    sym_mul_vec(m, bj[c], mv);
    for (int rr = 0; rr < 2; ++rr)
      r[rr][c] = bi[rr][0] * mv[0] + bi[rr][1] * mv[1] + bi[rr][2] * mv[2];
  }
}

SS_HD void inv2(const double m[2][2], double o[2][2]) {
  double det = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  if (fabs(det) < kTiny) det = kTiny;
  o[0][0] = m[1][1] / det;
  o[0][1] = -m[0][1] / det;
  o[1][0] = -m[1][0] / det;
  o[1][1] = m[0][0] / det;
}

// FAST (mixed tiers): u normalised by the refined reciprocal square root
// (any orthonormal basis of the plane serves the residual projection).
template <bool FAST = false>
SS_HD void plane_basis(const double n[3], double b[2][3]) {
  double a[3] = {0.0, 0.0, 0.0};
  if (fabs(n[2]) < 0.9)
    a[2] = 1.0;
  else
    a[0] = 1.0;
  double u[3] = {a[1] * n[2] - a[2] * n[1], a[2] * n[0] - a[0] * n[2],
                 a[0] * n[1] - a[1] * n[0]};
  if constexpr (FAST) {
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double iu = uu > kTiny * kTiny ? rsq64(uu) : 1.0 / kTiny;
    for (int k = 0; k < 3; ++k) u[k] *= iu;
  } else {
    double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    un = un > kTiny ? un : kTiny;
    for (int k = 0; k < 3; ++k) u[k] /= un;
  }
  b[0][0] = u[0];
  b[0][1] = u[1];
  b[0][2] = u[2];
  b[1][0] = n[1] * u[2] - n[2] * u[1];
  b[1][1] = n[2] * u[0] - n[0] * u[2];
  b[1][2] = n[0] * u[1] - n[1] * u[0];
}

template <int NMAX>
struct JointWork {
  double o[NMAX][3];      // scaled plane points
  double nr[NMAX][3];     // unit normals
  double B[NMAX][2][3];   // in-plane bases
  double ab[NMAX][2];
  double delta[NMAX][2];
  double dinv[NMAX][2][2];
  double y[NMAX][2];
  double A[NMAX][2][2];
  double Cu[NMAX][2][2];
  double f[NMAX][2];
};

// Loops over the stack's interfaces (i < N) and segments (k <= N) with the
// compile-time bound NMAX: shallow instantiations (NMAX <= 4) unroll them,
// so the per-lane state lives in registers instead of private memory.
template <int NMAX>
struct JUnroll {
  static constexpr int kU = NMAX <= 4 ? NMAX + 1 : 1;
};
#define SS_JLOOP(i)                                  \
  _Pragma("unroll JUnroll<NMAX>::kU") for (int i = 0; i < NMAX; ++i) if (i < N)
#define SS_JLOOP_SEG(k)                              \
  _Pragma("unroll JUnroll<NMAX>::kU") for (int k = 0; k <= NMAX; ++k) if (k <= N)
#define SS_JLOOP_BACK(i)                             \
  _Pragma("unroll JUnroll<NMAX>::kU") for (int i = NMAX - 1; i >= 0; --i) if (i < N)

// merit |F|^2 of offsets ab (+ lam * delta) and optionally F itself.
template <int NMAX>
SS_HD double joint_geom(JointWork<NMAX>& w, int N, const double ps[3], const double qs[3],
                        const double* nseg, double lam, bool with_delta, bool keep_f,
                        double s_out[NMAX + 1], double dh_out[NMAX + 1][3]) {
  double prev[3] = {ps[0], ps[1], ps[2]};
  double gprev[3] = {0, 0, 0};
  double merit = 0.0;
  SS_JLOOP_SEG(k) {
    double nxt[3];
    if (k < N) {
      double a0 = w.ab[k][0] + (with_delta ? lam * w.delta[k][0] : 0.0);
      double a1 = w.ab[k][1] + (with_delta ? lam * w.delta[k][1] : 0.0);
      for (int c = 0; c < 3; ++c) nxt[c] = w.o[k][c] + (w.B[k][0][c] * a0 + w.B[k][1][c] * a1);
    } else {
      nxt[0] = qs[0];
      nxt[1] = qs[1];
      nxt[2] = qs[2];
    }
    double d[3] = {nxt[0] - prev[0], nxt[1] - prev[1], nxt[2] - prev[2]};
    double s = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    s = s > kTiny ? s : kTiny;
    double dh[3] = {d[0] / s, d[1] / s, d[2] / s};
    double g[3] = {nseg[k] * dh[0], nseg[k] * dh[1], nseg[k] * dh[2]};
    if (s_out) {
      s_out[k] = s;
      dh_out[k][0] = dh[0];
      dh_out[k][1] = dh[1];
      dh_out[k][2] = dh[2];
    }
    if (k > 0) {
      int i = k - 1;
      double dg[3] = {gprev[0] - g[0], gprev[1] - g[1], gprev[2] - g[2]};
      double f0 = w.B[i][0][0] * dg[0] + w.B[i][0][1] * dg[1] + w.B[i][0][2] * dg[2];
      double f1 = w.B[i][1][0] * dg[0] + w.B[i][1][1] * dg[1] + w.B[i][1][2] * dg[2];
      merit += f0 * f0 + f1 * f1;
      if (keep_f) {
        w.f[i][0] = f0;
        w.f[i][1] = f1;
      }
    }
    gprev[0] = g[0];
    gprev[1] = g[1];
    gprev[2] = g[2];
    prev[0] = nxt[0];
    prev[1] = nxt[1];
    prev[2] = nxt[2];
  }
  return merit;
}

// Joint solve + polyline.  pts/nrs: local planes (unscaled); nseg: N+1
// media indices.  Outputs the crossings x (unscaled), cos(theta1/theta2) per
// interface, segment lengths s (N+1) and validity.
template <int NMAX>
SS_HD bool joint_crossings(const double p[3], const double q[3], int N,
                           const double (*pts)[3], const double (*nrs)[3],
                           const double* nseg, int n_newton, int n_backtrack,
                           double x[NMAX][3], double th1[NMAX], double th2[NMAX],
                           double s[NMAX + 1]) {
  JointWork<NMAX> w;
  // initial path: sequential chain over the local planes (25 iterations)
  double cur[3] = {p[0], p[1], p[2]};
  SS_JLOOP(i) {
    Snell64 sn;
    snell_fp64(cur, q, pts[i], nrs[i], nseg[i], nseg[i + 1], kLocalIters, sn);
    x[i][0] = sn.x[0];
    x[i][1] = sn.x[1];
    x[i][2] = sn.x[2];
    cur[0] = sn.x[0];
    cur[1] = sn.x[1];
    cur[2] = sn.x[2];
  }
  double dpq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  double scale = sqrt(dpq[0] * dpq[0] + dpq[1] * dpq[1] + dpq[2] * dpq[2]);
  scale = scale > kTiny ? scale : kTiny;
  double inv = 1.0 / scale;
  double ps[3] = {p[0] * inv, p[1] * inv, p[2] * inv};
  double qs[3] = {q[0] * inv, q[1] * inv, q[2] * inv};
  SS_JLOOP(i) {
    for (int c = 0; c < 3; ++c) {
      w.o[i][c] = pts[i][c] * inv;
      w.nr[i][c] = nrs[i][c];
    }
    plane_basis(w.nr[i], w.B[i]);
    double r[3] = {x[i][0] * inv - w.o[i][0], x[i][1] * inv - w.o[i][1],
                   x[i][2] * inv - w.o[i][2]};
    w.ab[i][0] = w.B[i][0][0] * r[0] + w.B[i][0][1] * r[1] + w.B[i][0][2] * r[2];
    w.ab[i][1] = w.B[i][1][0] * r[0] + w.B[i][1][1] * r[1] + w.B[i][1][2] * r[2];
  }
  const double mu = 1.8189894035458565e-12;  // eps^0.75 (float64)
  double sk[NMAX + 1], dhk[NMAX + 1][3];
  for (int it = 0; it < n_newton; ++it) {
    double m0 = joint_geom<NMAX>(w, N, ps, qs, nseg, 0.0, false, true, sk, dhk);
    // converged to the float64 floor: the remaining steps of the reference's
    // fixed budget move the offsets by rounding only (as flat_newton64)
    if (!(m0 > 1e-30)) break;
    // blocks
    double wprev[6];
    mat3_sym_outer(dhk[0], nseg[0] / sk[0], wprev);
    SS_JLOOP(i) {
      double wn[6], ws[6];
      mat3_sym_outer(dhk[i + 1], nseg[i + 1] / sk[i + 1], wn);
      for (int e = 0; e < 6; ++e) ws[e] = wprev[e] + wn[e];
      btmb(w.B[i], ws, w.B[i], w.A[i]);
      w.A[i][0][0] += mu;
      w.A[i][1][1] += mu;
      if (i < N - 1) {
        double cblk[2][2];
        btmb(w.B[i], wn, w.B[i + 1], cblk);
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) w.Cu[i][a][b] = -cblk[a][b];
      }
      for (int e = 0; e < 6; ++e) wprev[e] = wn[e];
    }
    // block Thomas: forward elimination
    SS_JLOOP(i) {
      double M[2][2] = {{w.A[i][0][0], w.A[i][0][1]}, {w.A[i][1][0], w.A[i][1][1]}};
      double yv[2] = {-w.f[i][0], -w.f[i][1]};
      if (i > 0) {
        // cpt = C_{i-1}^T dinv_{i-1}
        double cpt[2][2];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            cpt[a][b] = w.Cu[i - 1][0][a] * w.dinv[i - 1][0][b] +
                        w.Cu[i - 1][1][a] * w.dinv[i - 1][1][b];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            M[a][b] -= cpt[a][0] * w.Cu[i - 1][0][b] + cpt[a][1] * w.Cu[i - 1][1][b];
        yv[0] -= cpt[0][0] * w.y[i - 1][0] + cpt[0][1] * w.y[i - 1][1];
        yv[1] -= cpt[1][0] * w.y[i - 1][0] + cpt[1][1] * w.y[i - 1][1];
      }
      inv2(M, w.dinv[i]);
      w.y[i][0] = yv[0];
      w.y[i][1] = yv[1];
    }
    // back substitution
    double dn[2] = {0.0, 0.0};
    SS_JLOOP_BACK(i) {
      double r0 = w.y[i][0], r1 = w.y[i][1];
      if (i < N - 1) {
        r0 -= w.Cu[i][0][0] * dn[0] + w.Cu[i][0][1] * dn[1];
        r1 -= w.Cu[i][1][0] * dn[0] + w.Cu[i][1][1] * dn[1];
      }
      w.delta[i][0] = w.dinv[i][0][0] * r0 + w.dinv[i][0][1] * r1;
      w.delta[i][1] = w.dinv[i][1][0] * r0 + w.dinv[i][1][1] * r1;
      dn[0] = w.delta[i][0];
      dn[1] = w.delta[i][1];
    }
    // step halving: first decreasing candidate
    double lam = 1.0, sel = 0.0;
    for (int j = 0; j < n_backtrack; ++j) {
      double mc = joint_geom<NMAX>(w, N, ps, qs, nseg, lam, true, false, nullptr, nullptr);
      if (mc < m0) {
        sel = lam;
        break;
      }
      lam *= 0.5;
    }
    SS_JLOOP(i) {
      w.ab[i][0] += sel * w.delta[i][0];
      w.ab[i][1] += sel * w.delta[i][1];
    }
  }
  // final crossings (unscaled) and the honest polyline geometry
  SS_JLOOP(i)
    for (int c = 0; c < 3; ++c)
      x[i][c] = (w.o[i][c] + (w.B[i][0][c] * w.ab[i][0] + w.B[i][1][c] * w.ab[i][1])) * scale;
  double prev[3] = {p[0], p[1], p[2]};
  double din[3], dout[3];
  double resid = 0.0;
  bool sides = true, evan = false, fin = true;
  double sdir[NMAX + 1][3];
  SS_JLOOP_SEG(k) {
    const double* nxt = k < N ? x[k] : q;
    double d[3] = {nxt[0] - prev[0], nxt[1] - prev[1], nxt[2] - prev[2]};
    double sl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    s[k] = sl;
    double sm = sl > kTiny ? sl : kTiny;
    sdir[k][0] = d[0] / sm;
    sdir[k][1] = d[1] / sm;
    sdir[k][2] = d[2] / sm;
    fin = fin && isfinite(sl);
    prev[0] = nxt[0];
    prev[1] = nxt[1];
    prev[2] = nxt[2];
  }
  const double eps64 = kEps64;
  SS_JLOOP(i) {
    const double* n = w.nr[i];
    for (int c = 0; c < 3; ++c) {
      din[c] = sdir[i][c];
      dout[c] = sdir[i + 1][c];
    }
    double ct1 = n[0] * din[0] + n[1] * din[1] + n[2] * din[2];
    double ct2 = n[0] * dout[0] + n[1] * dout[1] + n[2] * dout[2];
    double t1[3] = {din[0] - ct1 * n[0], din[1] - ct1 * n[1], din[2] - ct1 * n[2]};
    double t2[3] = {dout[0] - ct2 * n[0], dout[1] - ct2 * n[1], dout[2] - ct2 * n[2]};
    double st1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    double st2 = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    th1[i] = atan2(st1, fabs(ct1));
    th2[i] = atan2(st2, fabs(ct2));
    double dG[3] = {nseg[i] * din[0] - nseg[i + 1] * dout[0],
                    nseg[i] * din[1] - nseg[i + 1] * dout[1],
                    nseg[i] * din[2] - nseg[i + 1] * dout[2]};
    double nd = n[0] * dG[0] + n[1] * dG[1] + n[2] * dG[2];
    double rt[3] = {dG[0] - nd * n[0], dG[1] - nd * n[1], dG[2] - nd * n[2]};
    double rn = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    resid = rn > resid ? rn : resid;
    fin = fin && isfinite(rn) && isfinite(x[i][0]) && isfinite(x[i][1]) && isfinite(x[i][2]);
    const double* pa = i == 0 ? p : x[i - 1];
    const double* pb = i == N - 1 ? q : x[i + 1];
    double hp = n[0] * (pa[0] - x[i][0]) + n[1] * (pa[1] - x[i][1]) + n[2] * (pa[2] - x[i][2]);
    double hn = n[0] * (pb[0] - x[i][0]) + n[1] * (pb[1] - x[i][1]) + n[2] * (pb[2] - x[i][2]);
    sides = sides && (hp * hn < 0.0);
    evan = evan || (nseg[i] * st1 > nseg[i + 1] * (1.0 + 64.0 * eps64));
  }
  bool converged = (resid <= 1e-9) && fin;
  return converged && sides && !evan && fin;
}

// Joint path (multilayer.py:408-457): mean-plane chain anchors -> cell
// lookups -> joint solve -> path sums.  N = padded stack depth; pads first.
template <int NMAX>
SS_HD void joint_path(const double p[3], const double q[3], const RefrConst& rc, int k_pad,
                      int n_newton, int n_backtrack, double cur[3], bool& valid,
                      double& opl, double& loss_db, double& sum_par, double& sum_perp,
                      double& tau2, double& c_first, double& c_last, double x_first[3]) {
  const int N = rc.n_crossed;
  const double* nl = rc.legs;
  const double* epsl = rc.legs + (N + 1);
  const double* attl = rc.legs + 2 * (N + 1);
  double pts[NMAX][3], nrs[NMAX][3];
  double c0[3] = {p[0], p[1], p[2]};
  SS_JLOOP(i) {
    const double* d = rc.desc + i * SS_IDESC_COUNT;
    double mp[3] = {d[SS_IDESC_MP], d[SS_IDESC_MP + 1], d[SS_IDESC_MP + 2]};
    double mn[3] = {d[SS_IDESC_MN], d[SS_IDESC_MN + 1], d[SS_IDESC_MN + 2]};
    Snell64 s;
    snell_fp64(c0, q, mp, mn, nl[i], nl[i + 1], kMeanIters, s);
    c0[0] = s.x[0];
    c0[1] = s.x[1];
    c0[2] = s.x[2];
    const double* cf = d + SS_IDESC_COEF;
    double rc0 = (s.x[0] * cf[0] + s.x[1] * cf[2]) + cf[4];
    double rc1 = (s.x[0] * cf[1] + s.x[1] * cf[3]) + cf[5];
    int row = clamp_round(rc0, (int)d[SS_IDESC_NY] - 1);
    int col = clamp_round(rc1, (int)d[SS_IDESC_NX] - 1);
    int nx = (int)d[SS_IDESC_NX];
    const double* g = rc.tables + ((long)d[SS_IDESC_TABOFF] + (long)row * nx + col) * 6;
    pts[i][0] = g[0];
    pts[i][1] = g[1];
    pts[i][2] = g[2];
    double nn = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    nrs[i][0] = g[3] / nn;
    nrs[i][1] = g[4] / nn;
    nrs[i][2] = g[5] / nn;
  }
  double x[NMAX][3], th1[NMAX], th2[NMAX], s[NMAX + 1];
  valid = joint_crossings<NMAX>(p, q, N, pts, nrs, nl, n_newton, n_backtrack, x, th1, th2, s);
  double c_surf = fmax(cos(th1[k_pad]), kCMin);
  tau2 = 1.0;
  opl = loss_db = sum_par = sum_perp = 0.0;
  SS_JLOOP(i) {
    double ci = fmax(cos(th1[i]), kCMin);
    double e1 = epsl[i], e2 = epsl[i + 1];
    double n1 = sqrt(e1), n2 = sqrt(e2);
    double sin2 = (e1 / e2) * (1.0 - ci * ci);
    double c2 = sqrt(fmax(1.0 - sin2, 0.0));
    double g = (n1 * ci - n2 * c2) / (n1 * ci + n2 * c2);
    tau2 *= (1.0 - g * g);
    if (rc.flags & SS_CROSS_ROUGH) {
      double kt = rc.k0 * (nl[i] * cos(th1[i]) - nl[i + 1] * cos(th2[i]));
      double sk = rc.desc[i * SS_IDESC_COUNT + SS_IDESC_SIGC] * kt;
      tau2 *= exp(-2.0 * (sk * sk));
    }
    double cp = i < k_pad ? c_surf : ci;
    opl += nl[i] * s[i];
    loss_db += s[i] * (attl[i] / 1000.0);
    sum_perp += s[i] / nl[i];
    sum_par += s[i] / (nl[i] * cp * cp);
  }
  c_first = c_surf;
  c_last = fmax(cos(th2[N - 1]), kCMin);
  cur[0] = x[N - 1][0];
  cur[1] = x[N - 1][1];
  cur[2] = x[N - 1][2];
  x_first[0] = x[0][0];
  x_first[1] = x[0][1];
  x_first[2] = x[0][2];
}

// Mixed tier of the joint path.  The crossings solve the same stationary
// problem as joint_path; only its root enters the result (the reference's
// damped Newton, started from the sequential chain over the local planes,
// reaches it within its 6-step budget whenever full Newton steps converge),
// so the search runs in float32 and one float64 step lands on the root:
//   pass 1 -- the mean-plane chain in float32 (coordinates relative to each
//     mean-plane point), the cell lookups from it; when any crossing lies
//     within kCellGuard of a rounding boundary the whole chain is taken in
//     float64 as the reference takes it;
//   initial path -- the sequential chain over the local planes, float32,
//     placed on the planes in float64;
//   Newton -- float32 on in-plane corrections to the initial path (segment
//     vectors = the initial path's, differenced in float64, plus the
//     corrections: thin layers keep their accuracy), coordinates relative
//     to q scaled by |p - q|, the
//     reference's block-tridiagonal system (Levenberg term included) and
//     step halving; a search where no step lowers |F|^2 above the float32
//     floor, or a damped one that took more than the reference's budget
//     less two steps, hands the pair to the float64 redo;
//   one float64 step -- F(ab) in float64, solved with the float32 block
//     factorisation (its 1e-7 relative error leaves |F| ~1e-14);
//   the polyline in float64 (validity, optical path) and the sums in float32.
// Returns false (float64 redo) unless the largest Snell residual after the
// float64 step is <= 1e-10 (a tenth of the reference's validity tolerance):
// then the float64 solve ends on the same root and the validity tests act
// on the same polyline.
template <int NMAX>
SS_HD bool joint_path_mixed(const double p[3], const double q[3], const RefrConst& rc,
                            PathSums& ps) {
  const int N = rc.n_crossed;
  const int k_pad = rc.k_pad;
  const double* nl = rc.legs;
  const double* epsl = rc.legs + (N + 1);
  const double* attl = rc.legs + 2 * (N + 1);
  double pts[NMAX][3], nrs[NMAX][3];
  // pass 1: mean-plane chain and cell lookups
  int row[NMAX], col[NMAX];
  bool exact = false;
  {
    double c0[3] = {p[0], p[1], p[2]};
    SS_JLOOP(i) {
      const double* d = rc.desc + i * SS_IDESC_COUNT;
      const double* cf = d + SS_IDESC_COEF;
      const double mp[3] = {d[SS_IDESC_MP], d[SS_IDESC_MP + 1], d[SS_IDESC_MP + 2]};
      float pr[3] = {(float)(c0[0] - mp[0]), (float)(c0[1] - mp[1]), (float)(c0[2] - mp[2])};
      float qr[3] = {(float)(q[0] - mp[0]), (float)(q[1] - mp[1]), (float)(q[2] - mp[2])};
      float mn[3] = {(float)d[SS_IDESC_MN], (float)d[SS_IDESC_MN + 1],
                     (float)d[SS_IDESC_MN + 2]};
      float xr[3];
      snell_anchor32(pr, qr, mn, (float)nl[i], (float)nl[i + 1], xr);
      const double x[3] = {(double)xr[0] + mp[0], (double)xr[1] + mp[1],
                           (double)xr[2] + mp[2]};
      const double rc0 = (x[0] * cf[0] + x[1] * cf[2]) + cf[4];
      const double rc1 = (x[0] * cf[1] + x[1] * cf[3]) + cf[5];
      const int ny = (int)d[SS_IDESC_NY], nx = (int)d[SS_IDESC_NX];
      if (ny > 1 || nx > 1) {
        const double f0 = rc0 - floor(rc0), f1 = rc1 - floor(rc1);
        exact = exact || fabs(f0 - 0.5) < kCellGuard || fabs(f1 - 0.5) < kCellGuard;
      }
      row[i] = clamp_round(rc0, ny - 1);
      col[i] = clamp_round(rc1, nx - 1);
      c0[0] = x[0];
      c0[1] = x[1];
      c0[2] = x[2];
    }
  }
  if (exact) {
    // near a lookup rounding boundary: the reference's float64 chain
    double c0[3] = {p[0], p[1], p[2]};
    SS_JLOOP(i) {
      const double* d = rc.desc + i * SS_IDESC_COUNT;
      const double mp[3] = {d[SS_IDESC_MP], d[SS_IDESC_MP + 1], d[SS_IDESC_MP + 2]};
      const double mn[3] = {d[SS_IDESC_MN], d[SS_IDESC_MN + 1], d[SS_IDESC_MN + 2]};
      Snell64 sn;
      snell_fp64(c0, q, mp, mn, nl[i], nl[i + 1], kMeanIters, sn);
      c0[0] = sn.x[0];
      c0[1] = sn.x[1];
      c0[2] = sn.x[2];
      const double* cf = d + SS_IDESC_COEF;
      double rc0, rc1;
      {
#pragma clang fp contract(off)
        rc0 = (sn.x[0] * cf[0] + sn.x[1] * cf[2]) + cf[4];
        rc1 = (sn.x[0] * cf[1] + sn.x[1] * cf[3]) + cf[5];
      }
      row[i] = clamp_round(rc0, (int)d[SS_IDESC_NY] - 1);
      col[i] = clamp_round(rc1, (int)d[SS_IDESC_NX] - 1);
    }
  }
  SS_JLOOP(i) {
    const double* d = rc.desc + i * SS_IDESC_COUNT;
    const int nx = (int)d[SS_IDESC_NX];
    const double* g = rc.tables + ((long)d[SS_IDESC_TABOFF] + (long)row[i] * nx + col[i]) * 6;
    pts[i][0] = g[0];
    pts[i][1] = g[1];
    pts[i][2] = g[2];
    const double inn = rsq64(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    nrs[i][0] = g[3] * inn;
    nrs[i][1] = g[4] * inn;
    nrs[i][2] = g[5] * inn;
  }
  // frame: relative to q, scaled by |p - q|
  const double dpq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  double scale = sqrt(dpq[0] * dpq[0] + dpq[1] * dpq[1] + dpq[2] * dpq[2]);
  scale = scale > kTiny ? scale : kTiny;
  const double inv = rcp64(scale);
  const double ps64[3] = {dpq[0] * inv, dpq[1] * inv, dpq[2] * inv};
  double B64[NMAX][2][3], x0[NMAX][3];
  float Bf[NMAX][2][3], ab[NMAX][2];
  float nsf[NMAX + 1];
  SS_JLOOP_SEG(k) { nsf[k] = (float)nl[k]; }
  // initial path: the chain over the local planes in float32 (relative to
  // each plane's point), placed on the planes in float64
  {
    float cur[3] = {(float)ps64[0], (float)ps64[1], (float)ps64[2]};
    SS_JLOOP(i) {
      plane_basis<true>(nrs[i], B64[i]);
      const float ni[3] = {(float)nrs[i][0], (float)nrs[i][1], (float)nrs[i][2]};
      double o64[3];
      float o[3];
      for (int c = 0; c < 3; ++c) {
        o64[c] = (pts[i][c] - q[c]) * inv;
        o[c] = (float)o64[c];
        Bf[i][0][c] = (float)B64[i][0][c];
        Bf[i][1][c] = (float)B64[i][1][c];
      }
      const float pr[3] = {cur[0] - o[0], cur[1] - o[1], cur[2] - o[2]};
      const float qr[3] = {-o[0], -o[1], -o[2]};
      float xr[3];
      snell_anchor32(pr, qr, ni, nsf[i], nsf[i + 1], xr);
      const float a0 = Bf[i][0][0] * xr[0] + Bf[i][0][1] * xr[1] + Bf[i][0][2] * xr[2];
      const float a1 = Bf[i][1][0] * xr[0] + Bf[i][1][1] * xr[1] + Bf[i][1][2] * xr[2];
      for (int c = 0; c < 3; ++c) x0[i][c] = o64[c] + (B64[i][0][c] * a0 + B64[i][1][c] * a1);
      cur[0] = (float)x0[i][0];
      cur[1] = (float)x0[i][1];
      cur[2] = (float)x0[i][2];
      ab[i][0] = 0.0f;  // the unknowns: in-plane corrections to x0
      ab[i][1] = 0.0f;
    }
  }
  // segment vectors of the initial path, differenced in float64 (thin
  // layers: a few metres against a path of tens of kilometres)
  float D[NMAX + 1][3];
  SS_JLOOP_SEG(k) {
    for (int c = 0; c < 3; ++c) {
      const double hi = k < N ? x0[k][c] : 0.0;
      const double lo = k > 0 ? x0[k - 1][c] : ps64[c];
      D[k][c] = (float)(hi - lo);
    }
  }
  // float32 Newton: F_i = B_i^T (n_i d_i - n_{i+1} d_{i+1}) on the polyline
  float f[NMAX][2], dinv[NMAX][2][2], Cu[NMAX][2][2], delta[NMAX][2];
  auto geom32 = [&](float lam, bool keep, float sk[NMAX + 1], float dk[NMAX + 1][3]) {
    float gprev[3] = {0.0f, 0.0f, 0.0f};
    float prevo[3] = {0.0f, 0.0f, 0.0f};  // in-plane correction of the last crossing
    float merit = 0.0f;
    SS_JLOOP_SEG(k) {
      float off[3] = {0.0f, 0.0f, 0.0f};
      if (k < N) {
        const float a0 = ab[k][0] + lam * delta[k][0];
        const float a1 = ab[k][1] + lam * delta[k][1];
        for (int c = 0; c < 3; ++c) off[c] = Bf[k][0][c] * a0 + Bf[k][1][c] * a1;
      }
      const float d[3] = {D[k][0] + (off[0] - prevo[0]), D[k][1] + (off[1] - prevo[1]),
                          D[k][2] + (off[2] - prevo[2])};
      const float s2 = fmaxf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], 1e-36f);
      const float is = f_rsq(s2);
      const float dh[3] = {d[0] * is, d[1] * is, d[2] * is};
      if (keep) {
        sk[k] = s2 * is;
        dk[k][0] = dh[0];
        dk[k][1] = dh[1];
        dk[k][2] = dh[2];
      }
      const float g[3] = {nsf[k] * dh[0], nsf[k] * dh[1], nsf[k] * dh[2]};
      if (k > 0) {
        const int i = k - 1;
        const float dg[3] = {gprev[0] - g[0], gprev[1] - g[1], gprev[2] - g[2]};
        const float f0 = Bf[i][0][0] * dg[0] + Bf[i][0][1] * dg[1] + Bf[i][0][2] * dg[2];
        const float f1 = Bf[i][1][0] * dg[0] + Bf[i][1][1] * dg[1] + Bf[i][1][2] * dg[2];
        merit += f0 * f0 + f1 * f1;
        if (keep) {
          f[i][0] = f0;
          f[i][1] = f1;
        }
      }
      gprev[0] = g[0];
      gprev[1] = g[1];
      gprev[2] = g[2];
      prevo[0] = off[0];
      prevo[1] = off[1];
      prevo[2] = off[2];
    }
    return merit;
  };
  SS_JLOOP(i) {
    delta[i][0] = 0.0f;
    delta[i][1] = 0.0f;
  }
  const float mu = 1.8189894035458565e-12f;
  bool factored = false;
  int steps = 0;
  bool damped = false;
  for (int it = 0; it < rc.n_newton + 2; ++it) {
    float sk[NMAX + 1], dk[NMAX + 1][3];
    const float m0 = geom32(0.0f, true, sk, dk);
    if (!(m0 == m0)) return false;
    // blocks and block Thomas (forward elimination)
    float wprev[6];
    {
      const float w0 = nsf[0] * f_rcp(sk[0]);
      wprev[0] = w0 * (1.0f - dk[0][0] * dk[0][0]);
      wprev[1] = w0 * (0.0f - dk[0][0] * dk[0][1]);
      wprev[2] = w0 * (0.0f - dk[0][0] * dk[0][2]);
      wprev[3] = w0 * (1.0f - dk[0][1] * dk[0][1]);
      wprev[4] = w0 * (0.0f - dk[0][1] * dk[0][2]);
      wprev[5] = w0 * (1.0f - dk[0][2] * dk[0][2]);
    }
    float yv[NMAX][2];
    SS_JLOOP(i) {
      float wn[6];
      const float w1 = nsf[i + 1] * f_rcp(sk[i + 1]);
      const float* dd = dk[i + 1];
      wn[0] = w1 * (1.0f - dd[0] * dd[0]);
      wn[1] = w1 * (0.0f - dd[0] * dd[1]);
      wn[2] = w1 * (0.0f - dd[0] * dd[2]);
      wn[3] = w1 * (1.0f - dd[1] * dd[1]);
      wn[4] = w1 * (0.0f - dd[1] * dd[2]);
      wn[5] = w1 * (1.0f - dd[2] * dd[2]);
      float A[2][2];
      for (int c = 0; c < 2; ++c) {
        const float* bj = Bf[i][c];
        const float mv[3] = {(wprev[0] + wn[0]) * bj[0] + (wprev[1] + wn[1]) * bj[1] +
                                 (wprev[2] + wn[2]) * bj[2],
                             (wprev[1] + wn[1]) * bj[0] + (wprev[3] + wn[3]) * bj[1] +
                                 (wprev[4] + wn[4]) * bj[2],
                             (wprev[2] + wn[2]) * bj[0] + (wprev[4] + wn[4]) * bj[1] +
                                 (wprev[5] + wn[5]) * bj[2]};
        for (int r = 0; r < 2; ++r)
          A[r][c] = Bf[i][r][0] * mv[0] + Bf[i][r][1] * mv[1] + Bf[i][r][2] * mv[2];
      }
      A[0][0] += mu;
      A[1][1] += mu;
      if (i < N - 1) {
        for (int c = 0; c < 2; ++c) {
          const float* bj = Bf[i + 1][c];
          const float mv[3] = {wn[0] * bj[0] + wn[1] * bj[1] + wn[2] * bj[2],
                               wn[1] * bj[0] + wn[3] * bj[1] + wn[4] * bj[2],
                               wn[2] * bj[0] + wn[4] * bj[1] + wn[5] * bj[2]};
          for (int r = 0; r < 2; ++r)
            Cu[i][r][c] = -(Bf[i][r][0] * mv[0] + Bf[i][r][1] * mv[1] + Bf[i][r][2] * mv[2]);
        }
      }
      for (int e = 0; e < 6; ++e) wprev[e] = wn[e];
      float y0 = -f[i][0], y1 = -f[i][1];
      if (i > 0) {
        float cpt[2][2];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            cpt[a][b] = Cu[i - 1][0][a] * dinv[i - 1][0][b] + Cu[i - 1][1][a] * dinv[i - 1][1][b];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            A[a][b] -= cpt[a][0] * Cu[i - 1][0][b] + cpt[a][1] * Cu[i - 1][1][b];
        y0 -= cpt[0][0] * yv[i - 1][0] + cpt[0][1] * yv[i - 1][1];
        y1 -= cpt[1][0] * yv[i - 1][0] + cpt[1][1] * yv[i - 1][1];
      }
      const float idet = f_rcp(A[0][0] * A[1][1] - A[0][1] * A[1][0]);
      dinv[i][0][0] = A[1][1] * idet;
      dinv[i][0][1] = -A[0][1] * idet;
      dinv[i][1][0] = -A[1][0] * idet;
      dinv[i][1][1] = A[0][0] * idet;
      yv[i][0] = y0;
      yv[i][1] = y1;
    }
    factored = true;
    if (!(m0 > 1e-13f)) break;  // at the float32 floor, factorised here
    float dn[2] = {0.0f, 0.0f};
    SS_JLOOP_BACK(i) {
      float r0 = yv[i][0], r1 = yv[i][1];
      if (i < N - 1) {
        r0 -= Cu[i][0][0] * dn[0] + Cu[i][0][1] * dn[1];
        r1 -= Cu[i][1][0] * dn[0] + Cu[i][1][1] * dn[1];
      }
      delta[i][0] = dinv[i][0][0] * r0 + dinv[i][0][1] * r1;
      delta[i][1] = dinv[i][1][0] * r0 + dinv[i][1][1] * r1;
      dn[0] = delta[i][0];
      dn[1] = delta[i][1];
    }
    // the reference's step halving: the first of 1, 1/2, ... lowering |F|^2
    float lam = 1.0f, sel = 0.0f;
    for (int b = 0; b < rc.n_backtrack; ++b, lam *= 0.5f) {
      if (geom32(lam, false, nullptr, nullptr) < m0) {
        sel = lam;
        break;
      }
    }
    if (sel == 0.0f) {
      if (m0 < 1e-10f) break;  // rounding-limited: the float64 step finishes
      return false;            // no step lowers |F|^2: the float64 solve decides
    }
    damped = damped || sel < 1.0f;
    ++steps;
    SS_JLOOP(i) {
      ab[i][0] += sel * delta[i][0];
      ab[i][1] += sel * delta[i][1];
      delta[i][0] = 0.0f;
      delta[i][1] = 0.0f;
    }
  }
  if (!factored) return false;
  // a damped search that needed most of the reference's step budget may
  // leave the float64 solve unconverged where this one converged
  if (damped && steps > rc.n_newton - 2) return false;
  // one float64 step: F(ab) in float64, the float32 factorisation
  double ab64[NMAX][2];
  SS_JLOOP(i) {
    ab64[i][0] = ab[i][0];
    ab64[i][1] = ab[i][1];
  }
  {
    double f64[NMAX][2];
    double prev[3] = {ps64[0], ps64[1], ps64[2]};
    double gprev[3] = {0.0, 0.0, 0.0};
    SS_JLOOP_SEG(k) {
      double nxt[3] = {0.0, 0.0, 0.0};
      if (k < N)
        for (int c = 0; c < 3; ++c)
          nxt[c] = x0[k][c] + (B64[k][0][c] * ab64[k][0] + B64[k][1][c] * ab64[k][1]);
      const double d[3] = {nxt[0] - prev[0], nxt[1] - prev[1], nxt[2] - prev[2]};
      const double is = rsq64(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double g[3] = {nl[k] * d[0] * is, nl[k] * d[1] * is, nl[k] * d[2] * is};
      if (k > 0) {
        const int i = k - 1;
        const double dg[3] = {gprev[0] - g[0], gprev[1] - g[1], gprev[2] - g[2]};
        f64[i][0] = B64[i][0][0] * dg[0] + B64[i][0][1] * dg[1] + B64[i][0][2] * dg[2];
        f64[i][1] = B64[i][1][0] * dg[0] + B64[i][1][1] * dg[1] + B64[i][1][2] * dg[2];
      }
      gprev[0] = g[0];
      gprev[1] = g[1];
      gprev[2] = g[2];
      prev[0] = nxt[0];
      prev[1] = nxt[1];
      prev[2] = nxt[2];
    }
    double y64[NMAX][2];
    SS_JLOOP(i) {
      double y0 = -f64[i][0], y1 = -f64[i][1];
      if (i > 0) {
        double cpt[2][2];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            cpt[a][b] = (double)Cu[i - 1][0][a] * dinv[i - 1][0][b] +
                        (double)Cu[i - 1][1][a] * dinv[i - 1][1][b];
        y0 -= cpt[0][0] * y64[i - 1][0] + cpt[0][1] * y64[i - 1][1];
        y1 -= cpt[1][0] * y64[i - 1][0] + cpt[1][1] * y64[i - 1][1];
      }
      y64[i][0] = y0;
      y64[i][1] = y1;
    }
    double dn[2] = {0.0, 0.0};
    SS_JLOOP_BACK(i) {
      double r0 = y64[i][0], r1 = y64[i][1];
      if (i < N - 1) {
        r0 -= (double)Cu[i][0][0] * dn[0] + (double)Cu[i][0][1] * dn[1];
        r1 -= (double)Cu[i][1][0] * dn[0] + (double)Cu[i][1][1] * dn[1];
      }
      dn[0] = (double)dinv[i][0][0] * r0 + (double)dinv[i][0][1] * r1;
      dn[1] = (double)dinv[i][1][0] * r0 + (double)dinv[i][1][1] * r1;
      ab64[i][0] += dn[0];
      ab64[i][1] += dn[1];
    }
  }
  // the polyline in float64
  double x[NMAX][3];
  SS_JLOOP(i) {
    for (int c = 0; c < 3; ++c)
      x[i][c] = q[c] + (x0[i][c] + (B64[i][0][c] * ab64[i][0] + B64[i][1][c] * ab64[i][1])) *
                           scale;
  }
  double s[NMAX + 1], sdir[NMAX + 1][3];
  bool fin = true;
  {
    double prev[3] = {p[0], p[1], p[2]};
    SS_JLOOP_SEG(k) {
      const double* nxt = k < N ? x[k] : q;
      const double d[3] = {nxt[0] - prev[0], nxt[1] - prev[1], nxt[2] - prev[2]};
      const double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double is = rsq64(s2 > kTiny * kTiny ? s2 : kTiny * kTiny);
      s[k] = s2 * is;
      sdir[k][0] = d[0] * is;
      sdir[k][1] = d[1] * is;
      sdir[k][2] = d[2] * is;
      fin = fin && isfinite(s[k]);
      prev[0] = nxt[0];
      prev[1] = nxt[1];
      prev[2] = nxt[2];
    }
  }
  double resid = 0.0;
  bool sides = true, evan = false;
  float ci_raw[NMAX], ct_raw[NMAX];
  SS_JLOOP(i) {
    const double* n = nrs[i];
    const double* din = sdir[i];
    const double* dout = sdir[i + 1];
    const double ct1 = n[0] * din[0] + n[1] * din[1] + n[2] * din[2];
    const double ct2 = n[0] * dout[0] + n[1] * dout[1] + n[2] * dout[2];
    const double t1[3] = {din[0] - ct1 * n[0], din[1] - ct1 * n[1], din[2] - ct1 * n[2]};
    const double st1 = sqrt64(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    ci_raw[i] = (float)fabs(ct1);
    ct_raw[i] = (float)fabs(ct2);
    const double dG[3] = {nl[i] * din[0] - nl[i + 1] * dout[0],
                          nl[i] * din[1] - nl[i + 1] * dout[1],
                          nl[i] * din[2] - nl[i + 1] * dout[2]};
    const double nd = n[0] * dG[0] + n[1] * dG[1] + n[2] * dG[2];
    const double rt[3] = {dG[0] - nd * n[0], dG[1] - nd * n[1], dG[2] - nd * n[2]};
    const double rn = sqrt64(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    resid = rn > resid ? rn : resid;
    fin = fin && isfinite(rn) && isfinite(x[i][0]) && isfinite(x[i][1]) && isfinite(x[i][2]);
    const double* pa = i == 0 ? p : x[i - 1];
    const double* pb = i == N - 1 ? q : x[i + 1];
    const double hp = n[0] * (pa[0] - x[i][0]) + n[1] * (pa[1] - x[i][1]) + n[2] * (pa[2] - x[i][2]);
    const double hn = n[0] * (pb[0] - x[i][0]) + n[1] * (pb[1] - x[i][1]) + n[2] * (pb[2] - x[i][2]);
    sides = sides && (hp * hn < 0.0);
    evan = evan || (nl[i] * st1 > nl[i + 1] * (1.0 + 64.0 * kEps64));
  }
  // converged to the root (the reference's own residual there is ~1e-15):
  // the validity tests below are the reference's on the same polyline
  if (!(resid <= 1e-10)) return false;
  // sums (float32) and the optical path (float64)
  const float c_surf = fmaxf(ci_raw[k_pad], (float)kCMin);
  float tau2 = 1.0f, loss = 0.0f, perp = 0.0f, par = 0.0f;
  double opl = 0.0;
  const bool rough = (rc.flags & SS_CROSS_ROUGH) != 0;
  SS_JLOOP(i) {
    const float ci = fmaxf(ci_raw[i], (float)kCMin);
    const float e1 = (float)epsl[i], e2 = (float)epsl[i + 1];
    const float n1 = (float)nl[i], n2 = (float)nl[i + 1];
    const float sin2 = (e1 * f_rcp(e2)) * (1.0f - ci * ci);
    const float c2 = f_sqrt(fmaxf(1.0f - sin2, 0.0f));
    const float g = (n1 * ci - n2 * c2) * f_rcp(n1 * ci + n2 * c2);
    tau2 *= 1.0f - g * g;
    if (rough) {
      const float kt = (float)rc.k0 * (n1 * ci_raw[i] - n2 * ct_raw[i]);
      const float sk = (float)rc.desc[i * SS_IDESC_COUNT + SS_IDESC_SIGC] * kt;
      tau2 *= f_exp(-2.0f * (sk * sk));
    }
    const float cp = i < k_pad ? c_surf : ci;
    const float sf = (float)s[i];
    opl += nl[i] * s[i];
    loss += sf * (float)(attl[i] / 1000.0);
    perp += sf * f_rcp(n1);
    par += sf * f_rcp(n1 * cp * cp);
  }
  ps.valid = fin && sides && !evan && resid <= 1e-9;
  ps.opl = opl;
  ps.loss_db = loss;
  ps.sum_par = par;
  ps.sum_perp = perp;
  ps.tau2 = tau2;
  ps.c_first = c_surf;
  ps.c_last = fmaxf(ct_raw[N - 1], (float)kCMin);
  ps.cur[0] = x[N - 1][0];
  ps.cur[1] = x[N - 1][1];
  ps.cur[2] = x[N - 1][2];
  ps.x_first[0] = x[0][0];
  ps.x_first[1] = x[0][1];
  ps.x_first[2] = x[0][2];
  return true;
}

// ------------------------------------------------------------ flat stack
// Joint solve specialised to a stack whose interfaces below the surface are
// horizontal planes.  The joint anchoring of such a plane cannot depend on
// where the chain crosses it, so only the surface is anchored: pass 1 of the
// joint solve runs the chain through the no-op pads (each a straight step
// toward q, so the last pad crossing lies on the segment p -> q) and then
// Snell against the surface mean plane (10 iterations) and the cell lookup,
// exactly as joint_path does.  The stationary path is then fixed by the
// horizontal slowness v, conserved below the surface: given v it is built
// upward from q (c_k = sqrt(n_k^2 - |v|^2); x_top = q - sum_{k>=2} d_k
// (v, -c_k)/c_k on the first horizontal plane; then up along u = (-v, c_1)
// to the anchored plane, s = x_top + lam u), and Snell at the surface,
// F(v) = B^T (n_0 (s - p)/|s - p| - (v, -c_1)) = 0, is solved by damped
// Newton with the joint solve's budgets and an analytic Jacobian, from the
// slowness of p -> s0 refracted at the anchored plane.  No inner solve: one
// sqrt and one division per layer per evaluation.  The pads lie on the
// straight air leg and change none of the sums (index-matched, booked with
// the surface cosine), so the air leg is booked as one segment.  Validity
// and sums follow joint_crossings / joint_path on the polyline; at the
// horizontal planes the tangential residual vanishes by construction (both
// sides carry the same v), so the residual test acts on the surface.

// z of the horizontal interface with absolute index i: its first table cell.
SS_HD double flat_z(const RefrConst& rc, int i) {
  return rc.tables[(long)rc.desc[i * SS_IDESC_COUNT + SS_IDESC_TABOFF] * 6 + 2];
}

struct FlatStack {
  const RefrConst* rc;
  const double* n;  // media indices from the air leg down (N + 1)
  int kp, N;        // pads, real interfaces (surface first)
  // height of real layer k (2..N) between its planes (the last ends at qz)
  SS_HD double d(int k, double qz) const {
    const double top = flat_z(*rc, kp + k - 1);
    const double bot = k == N ? qz : flat_z(*rc, kp + k);
    return top - bot;
  }
};

// The path of slowness v: F(v) and (jac) dF/dv in the plane basis B, the
// surface crossing s, its climb length factor lam, c_1, n.u; returns
// feasibility (every c real, lam > 0, the climbing ray meets the plane from
// below).
SS_HD bool flat_eval(const FlatStack& fs, const double p[3], const double q[3],
                     const double pt0[3], const double nr0[3], const double B[2][3],
                     const double v[2], bool jac, double F[2], double J[2][2], double s[3],
                     double& lam, double& c1, double& nu) {
  const double* nl = fs.n;
  const double vv = v[0] * v[0] + v[1] * v[1];
  bool ok = true;
  double a = 0.0, b = 0.0;  // sum d/c, sum d/c^3 over layers 2..N
  double xt[3] = {q[0], q[1], q[2]};
  if (fs.N > 1) {
    for (int k = 2; k <= fs.N; ++k) {
      const double ck2 = nl[k] * nl[k] - vv;
      ok = ok && (ck2 > 0.0);
      const double rk = 1.0 / sqrt(ck2);
      const double dk = fs.d(k, q[2]);
      a += dk * rk;
      b += dk * (rk * rk * rk);
    }
    xt[0] = q[0] - a * v[0];
    xt[1] = q[1] - a * v[1];
    xt[2] = flat_z(*fs.rc, fs.kp + 1);
  }
  const double c12 = nl[1] * nl[1] - vv;
  ok = ok && (c12 > 0.0);
  c1 = sqrt(c12);
  const double u[3] = {-v[0], -v[1], c1};
  nu = nr0[0] * u[0] + nr0[1] * u[1] + nr0[2] * u[2];
  lam = (nr0[0] * (pt0[0] - xt[0]) + nr0[1] * (pt0[1] - xt[1]) + nr0[2] * (pt0[2] - xt[2])) / nu;
  ok = ok && (nu > 0.0) && (lam > 0.0);
  for (int c = 0; c < 3; ++c) s[c] = xt[c] + lam * u[c];
  const double ds[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
  double L = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  L = L > kTiny ? L : kTiny;
  const double dh[3] = {ds[0] / L, ds[1] / L, ds[2] / L};
  const double n0 = nl[0];
  const double g[3] = {n0 * dh[0] - v[0], n0 * dh[1] - v[1], n0 * dh[2] + c1};
  F[0] = B[0][0] * g[0] + B[0][1] * g[1] + B[0][2] * g[2];
  F[1] = B[1][0] * g[0] + B[1][1] * g[1] + B[1][2] * g[2];
  if (jac) {
    for (int j = 0; j < 2; ++j) {
      // dx_top/dv_j = -(a e_j + b v v_j); du/dv_j = (-e_j, -v_j / c_1)
      const double dxt[3] = {-(a * (j == 0 ? 1.0 : 0.0) + b * v[0] * v[j]),
                             -(a * (j == 1 ? 1.0 : 0.0) + b * v[1] * v[j]), 0.0};
      const double du[3] = {j == 0 ? -1.0 : 0.0, j == 1 ? -1.0 : 0.0, -v[j] / c1};
      const double dl = -((nr0[0] * dxt[0] + nr0[1] * dxt[1] + nr0[2] * dxt[2]) +
                          lam * (nr0[0] * du[0] + nr0[1] * du[1] + nr0[2] * du[2])) / nu;
      double dsv[3];
      for (int c = 0; c < 3; ++c) dsv[c] = dxt[c] + u[c] * dl + lam * du[c];
      const double hd = dh[0] * dsv[0] + dh[1] * dsv[1] + dh[2] * dsv[2];
      // d(n0 dh)/dv_j - dk_1/dv_j, k_1 = (v, -c_1)
      const double G[3] = {n0 * (dsv[0] - dh[0] * hd) / L - (j == 0 ? 1.0 : 0.0),
                           n0 * (dsv[1] - dh[1] * hd) / L - (j == 1 ? 1.0 : 0.0),
                           n0 * (dsv[2] - dh[2] * hd) / L - v[j] / c1};
      J[0][j] = B[0][0] * G[0] + B[0][1] * G[1] + B[0][2] * G[2];
      J[1][j] = B[1][0] * G[0] + B[1][1] * G[1] + B[1][2] * G[2];
    }
  }
  return ok;
}

// Pass 1: the pads step straight toward q, then Snell against the surface
// mean plane and the cell lookup; the anchored plane (pt0, nr0) and the
// crossing projected onto it (s0).  mixed: float32 anchor in mean-plane
// relative coordinates, the float64 one only near a lookup rounding edge
// (as the sequential chain's mixed tier).
SS_HD void flat_anchor(const double p[3], const double q[3], const RefrConst& rc,
                       const FlatStack& fs, bool mixed, double pt0[3], double nr0[3],
                       double s0[3]) {
  const int kp = fs.kp;
  const double* nl = fs.n;
  double c0[3] = {p[0], p[1], p[2]};
  if (kp > 0) {
    const double zp = rc.desc[(kp - 1) * SS_IDESC_COUNT + SS_IDESC_MP + 2];
    const double f = mixed ? (p[2] - zp) * rcp64(p[2] - q[2]) : (p[2] - zp) / (p[2] - q[2]);
    for (int c = 0; c < 3; ++c) c0[c] = p[c] + f * (q[c] - p[c]);
    c0[2] = zp;
  }
  const double* d0 = rc.desc + kp * SS_IDESC_COUNT;
  const double* cf = d0 + SS_IDESC_COEF;
  double mp[3] = {d0[SS_IDESC_MP], d0[SS_IDESC_MP + 1], d0[SS_IDESC_MP + 2]};
  double mn[3] = {d0[SS_IDESC_MN], d0[SS_IDESC_MN + 1], d0[SS_IDESC_MN + 2]};
  double x[3];
  double rc0 = 0.0, rc1 = 0.0;
  bool exact = !mixed;
  if (mixed) {
    float pr[3] = {(float)(c0[0] - mp[0]), (float)(c0[1] - mp[1]), (float)(c0[2] - mp[2])};
    float qr[3] = {(float)(q[0] - mp[0]), (float)(q[1] - mp[1]), (float)(q[2] - mp[2])};
    float mf[3] = {(float)mn[0], (float)mn[1], (float)mn[2]};
    float xr[3];
    snell_anchor32(pr, qr, mf, (float)nl[0], (float)nl[1], xr);
    for (int c = 0; c < 3; ++c) x[c] = (double)xr[c] + mp[c];
    rc0 = (x[0] * cf[0] + x[1] * cf[2]) + cf[4];
    rc1 = (x[0] * cf[1] + x[1] * cf[3]) + cf[5];
    const double f0 = rc0 - floor(rc0), f1 = rc1 - floor(rc1);
    exact = fabs(f0 - 0.5) < kCellGuard || fabs(f1 - 0.5) < kCellGuard;
  }
  if (exact) {
    Snell64 sn;
    snell_fp64(c0, q, mp, mn, nl[0], nl[1], kMeanIters, sn);
    for (int c = 0; c < 3; ++c) x[c] = sn.x[c];
    {
#pragma clang fp contract(off)
      rc0 = (x[0] * cf[0] + x[1] * cf[2]) + cf[4];
      rc1 = (x[0] * cf[1] + x[1] * cf[3]) + cf[5];
    }
  }
  int row = clamp_round(rc0, (int)d0[SS_IDESC_NY] - 1);
  int col = clamp_round(rc1, (int)d0[SS_IDESC_NX] - 1);
  const double* g = rc.tables + ((long)d0[SS_IDESC_TABOFF] + (long)row * (int)d0[SS_IDESC_NX] +
                                 col) * 6;
  if (mixed) {
    // the refined reciprocal square root, as the mixed sequential chain
    const double inn = rsq64(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    for (int c = 0; c < 3; ++c) {
      pt0[c] = g[c];
      nr0[c] = g[3 + c] * inn;
    }
  } else {
    double nn = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    for (int c = 0; c < 3; ++c) {
      pt0[c] = g[c];
      nr0[c] = g[3 + c] / nn;
    }
  }
  double h = (x[0] - pt0[0]) * nr0[0] + (x[1] - pt0[1]) * nr0[1] + (x[2] - pt0[2]) * nr0[2];
  for (int c = 0; c < 3; ++c) s0[c] = x[c] - h * nr0[c];
}

// Start slowness: the ray p -> s0 refracted n0 -> n1 at the plane nr0.
// FAST (mixed tier): refined reciprocal / reciprocal square root estimates
// for the divisions and square roots (a start point only).
template <bool FAST = false>
SS_HD void flat_start(const double p[3], const double s0[3], const double nr0[3],
                      const double* nl, double v[2]) {
  const double ds[3] = {s0[0] - p[0], s0[1] - p[1], s0[2] - p[2]};
  double i0, i1, i2, r, cost;
  if constexpr (FAST) {
    const double iL = rsq64(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
    i0 = ds[0] * iL, i1 = ds[1] * iL, i2 = ds[2] * iL;
    r = nl[0] * rcp64(nl[1]);
  } else {
    const double L = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
    i0 = ds[0] / L, i1 = ds[1] / L, i2 = ds[2] / L;
    r = nl[0] / nl[1];
  }
  const double cosi = -(nr0[0] * i0 + nr0[1] * i1 + nr0[2] * i2);
  const double ct2 = 1.0 - r * r * (1.0 - cosi * cosi);
  cost = FAST ? sqrt64(ct2) : sqrt(ct2);
  const double m = nl[0] * cosi - nl[1] * cost;
  v[0] = nl[0] * i0 + m * nr0[0];
  v[1] = nl[0] * i1 + m * nr0[1];
}

// Layer constants of a flat stack for the mixed tier, staged in LDS once per
// workgroup (index k: real medium k, 0 the air above the surface; layers
// 2..N-1 lie between horizontal planes, layer N ends at each pair's target).
struct FlatLds {
  double n2[SS_MAX_CROSSED + 1];   // n_k^2
  double d[SS_MAX_CROSSED + 1];    // d_k = z_{k-1} - z_k, layers 2..N-1
  double dn2[SS_MAX_CROSSED + 1];  // |d_k| n_k^2: the layer's optical path is dn2 / c_k
  float n2f[SS_MAX_CROSSED + 1];
  float df[SS_MAX_CROSSED + 1];    // d_k
  float dnf[SS_MAX_CROSSED + 1];   // |d_k| n_k: segment length dnf / c_k
  float dn2f[SS_MAX_CROSSED + 1];  // |d_k| n_k^2
  float dabs[SS_MAX_CROSSED + 1];  // |d_k|
  float attf[SS_MAX_CROSSED + 1];  // attenuation dB per m of medium k
  float sigf[SS_MAX_CROSSED + 1];  // crossing roughness of interface k (k < N)
  double z1, zlast;                // planes kp + 1 and kp + N - 1
  double min_n2;                   // min n_k^2 over k = 2..N
  double min_n_evan;               // min n_k (1 + 64 eps) over k = 2..N
  float min_n2f;                   // min of n2f over k = 2..N
  int dpos;                        // d_k > 0 for every k = 2..N-1
};

// Thread k of the workgroup fills entry k (k <= N); stage_flat_mins (one
// thread, after a barrier) derives the stack-wide minima.
SS_HD void stage_flat_layer(const RefrConst& rc, FlatLds& L, int k) {
  const int kp = rc.k_pad, N = rc.n_crossed - rc.k_pad;
  if (k > N) return;
  const int NA = rc.n_crossed;
  const double n = rc.legs[kp + k];
  const double att = rc.legs[2 * (NA + 1) + kp + k];
  double d = 0.0;
  if (k >= 2 && k < N) d = flat_z(rc, kp + k - 1) - flat_z(rc, kp + k);
  const double ad = fabs(d);
  L.n2[k] = n * n;
  L.d[k] = d;
  L.dn2[k] = ad * (n * n);
  L.n2f[k] = (float)(n * n);
  L.df[k] = (float)d;
  L.dnf[k] = (float)(ad * n);
  L.dn2f[k] = (float)(ad * (n * n));
  L.dabs[k] = (float)ad;
  L.attf[k] = (float)(att / 1000.0);
  L.sigf[k] = k < N ? (float)rc.desc[(kp + k) * SS_IDESC_COUNT + SS_IDESC_SIGC] : 0.0f;
}

SS_HD void stage_flat_mins(const RefrConst& rc, FlatLds& L) {
  const int kp = rc.k_pad, N = rc.n_crossed - rc.k_pad;
  double mn2 = 1e300, mne = 1e300;
  float mnf = 3.0e38f;
  int dpos = 1;
  for (int k = 2; k <= N; ++k) {
    mn2 = fmin(mn2, L.n2[k]);
    mne = fmin(mne, rc.legs[kp + k] * (1.0 + 64.0 * kEps64));
    mnf = fminf(mnf, L.n2f[k]);
    if (k < N) dpos = dpos && (L.d[k] > 0.0);
  }
  L.min_n2 = mn2;
  L.min_n_evan = mne;
  L.min_n2f = mnf;
  L.dpos = dpos;
  L.z1 = N > 1 ? flat_z(rc, kp + 1) : 0.0;
  L.zlast = N > 1 ? flat_z(rc, kp + N - 1) : 0.0;
}

// Float32 twin of flat_eval (F and J only) in coordinates relative to s0:
// P, Q, O = p, q, pt0 minus s0; zt = first plane's z minus s0's; dN = the
// last layer's height (z_{N-1} - q_z); layer constants from L.
// Returns 0 feasible, 1 clearly infeasible (some c^2 < -1e-4, n.u < -1e-4
// or the plane more than 5 cm past the climbing ray's start: the float64
// test agrees), 2 infeasible within those margins.  Every c_k^2 = n_k^2 -
// |v|^2 is positive exactly when the smallest is (float subtraction is
// monotonic), so the layer loop carries no tests.
SS_HD int flat_eval32(const FlatStack& fs, const FlatLds& L, const float P[3], const float Q[3],
                      const float O[3], const float nr[3], const float B[2][3], float zt,
                      float dN, const float v[2], float F[2], float J[2][2]) {
  const float vv = v[0] * v[0] + v[1] * v[1];
  bool ok = true, clear = false;
  float a = 0.0f, b = 0.0f;
  float xt[3] = {Q[0], Q[1], Q[2]};
  if (fs.N > 1) {
    const float cmin = L.min_n2f - vv;
    ok = cmin > 0.0f;
    clear = cmin < -1e-4f;
    for (int k = 2; k < fs.N; ++k) {
      const float rk = f_rsq(fmaxf(L.n2f[k] - vv, 1e-30f));
      a += L.df[k] * rk;
      b += L.df[k] * (rk * rk * rk);
    }
    const float rN = f_rsq(fmaxf(L.n2f[fs.N] - vv, 1e-30f));
    a += dN * rN;
    b += dN * (rN * rN * rN);
    xt[0] = Q[0] - a * v[0];
    xt[1] = Q[1] - a * v[1];
    xt[2] = zt;
  }
  const float c12 = L.n2f[1] - vv;
  ok = ok && (c12 > 0.0f);
  clear = clear || (c12 < -1e-4f);
  const float c1 = f_sqrt(fmaxf(c12, 1e-30f));
  const float ic1 = f_rcp(c1);
  const float u[3] = {-v[0], -v[1], c1};
  const float nu = nr[0] * u[0] + nr[1] * u[1] + nr[2] * u[2];
  const float inu = f_rcp(nu);
  const float num = nr[0] * (O[0] - xt[0]) + nr[1] * (O[1] - xt[1]) + nr[2] * (O[2] - xt[2]);
  const float lam = num * inu;
  ok = ok && (nu > 0.0f) && (lam > 0.0f);
  clear = clear || (nu < -1e-4f) || (nu > 1e-4f && num < -0.05f);
  float sv[3];
  for (int c = 0; c < 3; ++c) sv[c] = xt[c] + lam * u[c];
  const float ds[3] = {sv[0] - P[0], sv[1] - P[1], sv[2] - P[2]};
  const float Ls = f_sqrt(fmaxf(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2], 1e-30f));
  const float iL = f_rcp(Ls);
  const float dh[3] = {ds[0] * iL, ds[1] * iL, ds[2] * iL};
  const float n0 = (float)fs.n[0];
  const float g[3] = {n0 * dh[0] - v[0], n0 * dh[1] - v[1], n0 * dh[2] + c1};
  F[0] = B[0][0] * g[0] + B[0][1] * g[1] + B[0][2] * g[2];
  F[1] = B[1][0] * g[0] + B[1][1] * g[1] + B[1][2] * g[2];
  for (int j = 0; j < 2; ++j) {
    const float dxt[3] = {-(a * (j == 0 ? 1.0f : 0.0f) + b * v[0] * v[j]),
                          -(a * (j == 1 ? 1.0f : 0.0f) + b * v[1] * v[j]), 0.0f};
    const float du[3] = {j == 0 ? -1.0f : 0.0f, j == 1 ? -1.0f : 0.0f, -v[j] * ic1};
    const float dl = -((nr[0] * dxt[0] + nr[1] * dxt[1] + nr[2] * dxt[2]) +
                       lam * (nr[0] * du[0] + nr[1] * du[1] + nr[2] * du[2])) * inu;
    float dsv[3];
    for (int c = 0; c < 3; ++c) dsv[c] = dxt[c] + u[c] * dl + lam * du[c];
    const float hd = dh[0] * dsv[0] + dh[1] * dsv[1] + dh[2] * dsv[2];
    const float G[3] = {n0 * (dsv[0] - dh[0] * hd) * iL - (j == 0 ? 1.0f : 0.0f),
                        n0 * (dsv[1] - dh[1] * hd) * iL - (j == 1 ? 1.0f : 0.0f),
                        n0 * (dsv[2] - dh[2] * hd) * iL - v[j] * ic1};
    J[0][j] = B[0][0] * G[0] + B[0][1] * G[1] + B[0][2] * G[2];
    J[1][j] = B[1][0] * G[0] + B[1][1] * G[1] + B[1][2] * G[2];
  }
  return ok ? 0 : (clear ? 1 : 2);
}

// Float64 damped Newton on v with the joint solve's budgets: of lam = 1,
// 1/2, ... (n_backtrack steps) the first feasible step that lowers |F|^2 is
// taken; none ends the solve.
SS_HD void flat_newton64(const FlatStack& fs, const double p[3], const double q[3],
                         const double pt0[3], const double nr0[3], const double B[2][3],
                         int n_newton, int n_backtrack, double v[2]) {
  double F[2], J[2][2], s[3], lam, c1, nu;
  flat_eval(fs, p, q, pt0, nr0, B, v, true, F, J, s, lam, c1, nu);
  for (int it = 0; it < n_newton; ++it) {
    const double m0 = F[0] * F[0] + F[1] * F[1];
    if (!(m0 > 1e-30)) break;
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double st0 = (J[1][1] * F[0] - J[0][1] * F[1]) / det;
    const double st1 = (J[0][0] * F[1] - J[1][0] * F[0]) / det;
    double lm = 1.0;
    bool taken = false;
    for (int bt = 0; bt < n_backtrack && !taken; ++bt, lm *= 0.5) {
      const double vt[2] = {v[0] - lm * st0, v[1] - lm * st1};
      double Ft[2], Jt[2][2], stt[3], lt, c1t, nut;
      const bool okt = flat_eval(fs, p, q, pt0, nr0, B, vt, true, Ft, Jt, stt, lt, c1t, nut);
      const double mt = Ft[0] * Ft[0] + Ft[1] * Ft[1];
      if (okt && isfinite(mt) && mt < m0) {
        v[0] = vt[0];
        v[1] = vt[1];
        F[0] = Ft[0];
        F[1] = Ft[1];
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) J[r][c] = Jt[r][c];
        taken = true;
      }
    }
    if (!taken) break;
  }
}

// The polyline of slowness v in one pass over the layers (float64): the
// joint path's validity and sums, the surface residual, cur and x_first.
// Segment i (medium i above interface i): the air leg |s - p|, layer 1 the
// climb |lam| n_1, layer k >= 2 |d_k| n_k / c_k (polyline lengths also where
// an unsolved lane left a negative height: the dropped power of invalid
// lanes is booked from them).  At a horizontal plane the incidence and
// transmission cosines are c_k/n_k and c_{k+1}/n_{k+1}, so the TE
// coefficient is (c_k - c_{k+1})/(c_k + c_{k+1}) (the reference's formula
// with its clamp where c_k/n_k < C_MIN).
SS_HD double flat_final(const FlatStack& fs, const double p[3], const double q[3],
                        const double pt0[3], const double nr0[3], const double v[2],
                        PathSums& ps) {
  const RefrConst& rc = *fs.rc;
  const int NA = rc.n_crossed, kp = fs.kp, N = fs.N;
  const double* nl = fs.n;
  const double* epsl = rc.legs + (NA + 1) + kp;
  const double* attl = rc.legs + 2 * (NA + 1) + kp;
  const bool rough = (rc.flags & SS_CROSS_ROUGH) != 0;
  const double vv = v[0] * v[0] + v[1] * v[1];
  const double vm = sqrt(vv);
  bool sides = true, evan = false, fin = true;
  double tau2 = 1.0, opl = 0.0, loss = 0.0, perp = 0.0, par = 0.0;
  // interface i from medium i to i + 1 with cosines ci_raw -> ct_raw
  auto cross = [&](int i, double ci_raw, double ct_raw, bool flat) {
    const double ci = fmax(ci_raw, kCMin);
    double gm;
    if (flat && ci_raw >= kCMin) {
      const double a0 = nl[i] * ci_raw, a1 = nl[i + 1] * ct_raw;   // c_i, c_{i+1}
      gm = (a0 - a1) / (a0 + a1);
    } else {
      const double e1 = epsl[i], e2 = epsl[i + 1];
      const double sin2 = (e1 / e2) * (1.0 - ci * ci);
      const double cc2 = sqrt(fmax(1.0 - sin2, 0.0));
      gm = (nl[i] * ci - nl[i + 1] * cc2) / (nl[i] * ci + nl[i + 1] * cc2);
    }
    tau2 *= (1.0 - gm * gm);
    if (rough) {
      const double kt = rc.k0 * (nl[i] * ci_raw - nl[i + 1] * ct_raw);
      const double sk = rc.desc[(kp + i) * SS_IDESC_COUNT + SS_IDESC_SIGC] * kt;
      tau2 *= exp(-2.0 * (sk * sk));
    }
    return ci;
  };
  const double c12 = nl[1] * nl[1] - vv;
  const double c1 = sqrt(c12);
  fin = fin && isfinite(c1);
  // layers 2..N: x_top offset, the segments above planes 2..N-1 and the
  // interfaces 1..N-1
  double a = 0.0, c_prev = c1, c_last = 0.0;
  for (int k = 2; k <= N; ++k) {
    const double ck = sqrt(nl[k] * nl[k] - vv);
    const double rk = 1.0 / ck;
    const double dk = fs.d(k, q[2]);
    a += dk * rk;
    fin = fin && isfinite(rk);
    sides = sides && (dk > 0.0);
    evan = evan || (vm > nl[k] * (1.0 + 64.0 * kEps64));
    // interface k-1 (medium k-1 -> k)
    cross(k - 1, c_prev / nl[k - 1], ck / nl[k], true);
    if (k < N) {
      // segment k: length |d_k| n_k / c_k; incidence cosine c_k / n_k
      const double ad = fabs(dk);
      const double len = ad * nl[k] * rk;
      const double cik = ck / nl[k];
      const double ci = fmax(cik, kCMin);
      opl += nl[k] * len;
      loss += len * (attl[k] / 1000.0);
      perp += ad * rk;
      par += len / (nl[k] * ci * ci);
    }
    c_prev = ck;
    c_last = ck;
  }
  double xt[3] = {q[0], q[1], q[2]};
  if (N > 1) {
    xt[0] = q[0] - a * v[0];
    xt[1] = q[1] - a * v[1];
    xt[2] = flat_z(rc, kp + 1);
  }
  const double u[3] = {-v[0], -v[1], c1};
  const double nu = nr0[0] * u[0] + nr0[1] * u[1] + nr0[2] * u[2];
  const double lam =
      (nr0[0] * (pt0[0] - xt[0]) + nr0[1] * (pt0[1] - xt[1]) + nr0[2] * (pt0[2] - xt[2])) / nu;
  double s[3];
  for (int c = 0; c < 3; ++c) s[c] = xt[c] + lam * u[c];
  fin = fin && isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]);
  // pads: straddled by p and s; x_first on the top pad
  for (int i = 0; i < kp; ++i) {
    const double zp = rc.desc[i * SS_IDESC_COUNT + SS_IDESC_MP + 2];
    sides = sides && ((p[2] - zp) * (s[2] - zp) < 0.0);
  }
  if (kp > 0) {
    const double zp = rc.desc[SS_IDESC_MP + 2];
    const double f = (p[2] - zp) / (p[2] - s[2]);
    for (int c = 0; c < 3; ++c) ps.x_first[c] = p[c] + f * (s[c] - p[c]);
    ps.x_first[2] = zp;
  } else {
    for (int c = 0; c < 3; ++c) ps.x_first[c] = s[c];
  }
  // surface: the honest residual from the air leg and the layer-1 direction
  const double ds[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
  const double sa = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  const double sam = sa > kTiny ? sa : kTiny;
  const double din[3] = {ds[0] / sam, ds[1] / sam, ds[2] / sam};
  const double dout[3] = {v[0] / nl[1], v[1] / nl[1], -c1 / nl[1]};
  const double ct1 = nr0[0] * din[0] + nr0[1] * din[1] + nr0[2] * din[2];
  const double ct2 = nr0[0] * dout[0] + nr0[1] * dout[1] + nr0[2] * dout[2];
  const double t1[3] = {din[0] - ct1 * nr0[0], din[1] - ct1 * nr0[1], din[2] - ct1 * nr0[2]};
  const double st1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  const double dG[3] = {nl[0] * din[0] - nl[1] * dout[0], nl[0] * din[1] - nl[1] * dout[1],
                        nl[0] * din[2] - nl[1] * dout[2]};
  const double nd = nr0[0] * dG[0] + nr0[1] * dG[1] + nr0[2] * dG[2];
  const double rt[3] = {dG[0] - nd * nr0[0], dG[1] - nd * nr0[1], dG[2] - nd * nr0[2]};
  const double resid = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
  fin = fin && isfinite(resid) && isfinite(sa);
  sides = sides &&
          ((nr0[0] * (p[0] - s[0]) + nr0[1] * (p[1] - s[1]) + nr0[2] * (p[2] - s[2])) *
               (-lam * nu) < 0.0);
  evan = evan || (nl[0] * st1 > nl[1] * (1.0 + 64.0 * kEps64));
  const double c_surf = cross(0, fabs(ct1), fabs(ct2), false);
  opl += nl[0] * sa;
  loss += sa * (attl[0] / 1000.0);
  perp += sa / nl[0];
  par += sa / (nl[0] * c_surf * c_surf);
  if (N > 1) {
    // layer 1: the climb
    const double len = fabs(lam) * nl[1];
    const double ci = fmax(c1 / nl[1], kCMin);
    opl += nl[1] * len;
    loss += len * (attl[1] / 1000.0);
    perp += len / nl[1];
    par += len / (nl[1] * ci * ci);
    const double dN = fs.d(N, q[2]);
    ps.cur[0] = q[0] - dN * v[0] / c_last;
    ps.cur[1] = q[1] - dN * v[1] / c_last;
    ps.cur[2] = flat_z(rc, kp + N - 1);
    ps.c_last = fmax(c_last / nl[N], kCMin);
  } else {
    for (int c = 0; c < 3; ++c) ps.cur[c] = s[c];
    ps.c_last = fmax(fabs(ct2), kCMin);
  }
  ps.valid = (resid <= 1e-9) && fin && sides && !evan;
  ps.opl = opl;
  ps.loss_db = loss;
  ps.sum_par = par;
  ps.sum_perp = perp;
  ps.tau2 = tau2;
  ps.c_first = c_surf;
  return resid;
}

SS_HD void flat_path(const double p[3], const double q[3], const RefrConst& rc,
                     PathSums& ps) {
  FlatStack fs;
  fs.rc = &rc;
  fs.kp = rc.k_pad;
  fs.N = rc.n_crossed - rc.k_pad;
  fs.n = rc.legs + rc.k_pad;
  double pt0[3], nr0[3], s0[3], B[2][3], v[2];
  flat_anchor(p, q, rc, fs, false, pt0, nr0, s0);
  plane_basis(nr0, B);
  flat_start(p, s0, nr0, fs.n, v);
  flat_newton64(fs, p, q, pt0, nr0, B, rc.n_newton, rc.n_backtrack, v);
  flat_final(fs, p, q, pt0, nr0, v, ps);
}

// Mixed tier: float32 anchor, float32 Newton in coordinates relative to s0,
// one float64 Newton step, then the polyline.  Returns false (float64 redo
// of the pair) unless the float32 phase took only full steps and the result
// is valid with a surface residual <= 1e-12 -- the cases where the float64
// solve reaches the same stationary point -- or the lane is stuck: no first
// step is feasible, its full step and every halving the float64 solve would
// try being clearly infeasible, so that solve stays at the start (the lane
// has no valid path: the anchored surface plane passes below the first
// horizontal plane) and the lane is booked from the start path.
//
// The float64 work is one layer loop at the float32 solution v: per layer
// one refined rsq64 gives 1/c_k, which sums a = sum d_k / c_k (the horizontal
// reach, hence F(v) and the crossings) and the layers' optical path
// sum |d_k| n_k^2 / c_k.  The Newton step dv = -J^-1 F(v) takes the float32
// Jacobian (its 1e-7 relative error leaves |F| ~ 1e-14 after the step, as the
// float64 one would), and the sums follow the step to first order: da/dv =
// b v, b = sum d_k / c_k^3, and likewise for the optical path and 1/c_N; the
// second-order remainder is ~b |dv|^2 ~ 1e-11 m against a bin guard of
// 1e-7 bins (2.5e-7 m at 60 MHz sampling).  Fresnel, crossing roughness,
// attenuation and the ray-tube sums are float32 from the same loop (at v;
// their change over dv is ~1e-7 relative).  The stack-wide tests (every
// c_k real, no evanescence, every layer straddled) compare |v| against the
// staged minima.
SS_HD bool flat_path_mixed(const double p[3], const double q[3], const RefrConst& rc,
                           const FlatLds& Ls, PathSums& ps) {
  FlatStack fs;
  fs.rc = &rc;
  fs.kp = rc.k_pad;
  fs.N = rc.n_crossed - rc.k_pad;
  fs.n = rc.legs + rc.k_pad;
  const int N = fs.N, kp = fs.kp;
  const double* nl = fs.n;
  double pt0[3], nr0[3], s0[3], B[2][3], v[2];
  flat_anchor(p, q, rc, fs, true, pt0, nr0, s0);
  plane_basis<true>(nr0, B);
  flat_start<true>(p, s0, nr0, nl, v);
  const double dN = N > 1 ? Ls.zlast - q[2] : 0.0;
  bool stuck = false;
  float Jf[2][2] = {{1.0f, 0.0f}, {0.0f, 1.0f}};
  {
    float P[3], Q[3], O[3], nr[3], Bf[2][3];
    for (int c = 0; c < 3; ++c) {
      P[c] = (float)(p[c] - s0[c]);
      Q[c] = (float)(q[c] - s0[c]);
      O[c] = (float)(pt0[c] - s0[c]);
      nr[c] = (float)nr0[c];
      Bf[0][c] = (float)B[0][c];
      Bf[1][c] = (float)B[1][c];
    }
    const float zt = N > 1 ? (float)(Ls.z1 - s0[2]) : Q[2];
    const float dNf = (float)dN;
    const float lm_min = ldexpf(1.0f, 1 - rc.n_backtrack);
    float vf[2] = {(float)v[0], (float)v[1]}, vt[2] = {vf[0], vf[1]};
    float F[2] = {0.0f, 0.0f}, st[2] = {0.0f, 0.0f};
    float m0 = 0.0f, lm = 1.0f;
    int steps = 0;
    bool base = true, halving = false, ok = false, fail = false;
    for (int ev = 0; ev < 40; ++ev) {
      float Ft[2], Jt[2][2];
      const int ft = flat_eval32(fs, Ls, P, Q, O, nr, Bf, zt, dNf, vt, Ft, Jt);
      const float mt = Ft[0] * Ft[0] + Ft[1] * Ft[1];
      bool take = false;
      if (base) {
        take = true;
        ok = ft == 0;
        base = false;
      } else if (halving) {
        // the float64 solve's halvings of the first step, all clearly infeasible?
        if (ft != 1) {
          fail = true;
          break;
        }
        if (lm <= lm_min) {
          stuck = true;
          break;
        }
        lm *= 0.5f;
        vt[0] = vf[0] - lm * st[0];
        vt[1] = vf[1] - lm * st[1];
        continue;
      } else if (ft == 0 && mt < m0) {
        take = true;
        ok = true;
        ++steps;
      } else if (!(m0 > 1e-10f)) {
        break;  // at the float32 floor the phase just ends
      } else if (steps == 0 && ft == 1) {
        if (rc.n_backtrack <= 1) {
          stuck = true;
          break;
        }
        halving = true;
        lm = 0.5f;
        vt[0] = vf[0] - lm * st[0];
        vt[1] = vf[1] - lm * st[1];
        continue;
      } else {
        fail = true;  // a damped step: the float64 solve decides
        break;
      }
      if (take) {
        vf[0] = vt[0];
        vf[1] = vt[1];
        F[0] = Ft[0];
        F[1] = Ft[1];
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) Jf[r][c] = Jt[r][c];
      }
      m0 = F[0] * F[0] + F[1] * F[1];
      if (!(m0 > 1e-13f) || steps == 4) break;
      const float idet = f_rcp(Jf[0][0] * Jf[1][1] - Jf[0][1] * Jf[1][0]);
      st[0] = (Jf[1][1] * F[0] - Jf[0][1] * F[1]) * idet;
      st[1] = (Jf[0][0] * F[1] - Jf[1][0] * F[0]) * idet;
      vt[0] = vf[0] - st[0];
      vt[1] = vf[1] - st[1];
    }
    if (fail || (!stuck && !ok)) return false;
    if (!stuck) {
      v[0] = vf[0];
      v[1] = vf[1];
    }
  }
  // the float64 layer loop at v, with the float32 radiometry sums
  const double vv = v[0] * v[0] + v[1] * v[1];
  const double c12 = Ls.n2[1] - vv;
  const double c1 = sqrt64(c12);
  const float c1f = (float)c1;
  const bool rough = (rc.flags & SS_CROSS_ROUGH) != 0;
  const float k0f = (float)rc.k0;
  double a = 0.0, o = 0.0, rN = 0.0;
  float b = 0.0f, o3 = 0.0f, tau2 = 1.0f, loss = 0.0f, perp = 0.0f, par = 0.0f;
  float c_prev = c1f, c_lastf = c1f;
  // interface i between cosines-times-index c_i -> c_{i+1} of two horizontal
  // media: TE coefficient (c_i - c_{i+1}) / (c_i + c_{i+1}), crossing
  // roughness exp(-2 (sigma k0 (c_i - c_{i+1}))^2)
  auto cross_flat = [&](int i, float ca, float cb) {
    const float gm = (ca - cb) * f_rcp(ca + cb);
    tau2 *= 1.0f - gm * gm;
    if (rough) {
      const float sk = Ls.sigf[i] * (k0f * (ca - cb));
      tau2 *= f_exp(-2.0f * (sk * sk));
    }
  };
  bool lay_ok = true;
  if (N > 1) {
    lay_ok = vv < Ls.min_n2;
    for (int k = 2; k < N; ++k) {
      const double rk = rsq64_1(Ls.n2[k] - vv);
      a = fma(Ls.d[k], rk, a);
      o = fma(Ls.dn2[k], rk, o);
      const float rf = (float)rk;
      const float r3 = rf * rf * rf;
      const float ck = f_rcp(rf);
      b += Ls.df[k] * r3;
      o3 += Ls.dn2f[k] * r3;
      cross_flat(k - 1, c_prev, ck);
      const float len = Ls.dnf[k] * rf;
      loss += len * Ls.attf[k];
      perp += Ls.dabs[k] * rf;
      c_prev = ck;
    }
    rN = rsq64(Ls.n2[N] - vv);
    a = fma(dN, rN, a);
    const float rf = (float)rN;
    b += (float)dN * (rf * rf * rf);
    c_lastf = f_rcp(rf);
    cross_flat(N - 1, c_prev, c_lastf);
  }
  par = o3;  // sum |d_k| n_k^2 / c_k^3 = sum len_k / (n_k cos_k^2)
  double u[3] = {-v[0], -v[1], c1};
  double xt[3] = {q[0], q[1], q[2]};
  if (N > 1) {
    xt[0] = q[0] - a * v[0];
    xt[1] = q[1] - a * v[1];
    xt[2] = Ls.z1;
  }
  double nu = nr0[0] * u[0] + nr0[1] * u[1] + nr0[2] * u[2];
  double lam = (nr0[0] * (pt0[0] - xt[0]) + nr0[1] * (pt0[1] - xt[1]) +
                nr0[2] * (pt0[2] - xt[2])) * rcp64(nu);
  double vn[2] = {v[0], v[1]}, c1n = c1;
  if (!stuck) {
    // feasibility at v, F(v), the step
    if (!(lay_ok && c12 > 0.0 && nu > 0.0 && lam > 0.0)) return false;
    double s[3];
    for (int c = 0; c < 3; ++c) s[c] = xt[c] + lam * u[c];
    const double ds[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
    const double iL = rsq64(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
    const double n0 = nl[0];
    const double g[3] = {n0 * ds[0] * iL - v[0], n0 * ds[1] * iL - v[1], n0 * ds[2] * iL + c1};
    const double F0 = B[0][0] * g[0] + B[0][1] * g[1] + B[0][2] * g[2];
    const double F1 = B[1][0] * g[0] + B[1][1] * g[1] + B[1][2] * g[2];
    const double j00 = Jf[0][0], j01 = Jf[0][1], j10 = Jf[1][0], j11 = Jf[1][1];
    const double idet = rcp64(j00 * j11 - j01 * j10);
    const double dv0 = -(j11 * F0 - j01 * F1) * idet;
    const double dv1 = -(j00 * F1 - j10 * F0) * idet;
    const double vdv = v[0] * dv0 + v[1] * dv1;
    vn[0] = v[0] + dv0;
    vn[1] = v[1] + dv1;
    if (N > 1) {
      a = fma((double)b, vdv, a);
      o = fma((double)o3, vdv, o);
      rN = fma(rN * rN * rN, vdv, rN);
    }
    const double vvn = vn[0] * vn[0] + vn[1] * vn[1];
    c1n = sqrt64(Ls.n2[1] - vvn);
    lay_ok = N > 1 ? vvn < Ls.min_n2 : true;
    u[0] = -vn[0];
    u[1] = -vn[1];
    u[2] = c1n;
    if (N > 1) {
      xt[0] = q[0] - a * vn[0];
      xt[1] = q[1] - a * vn[1];
    }
    nu = nr0[0] * u[0] + nr0[1] * u[1] + nr0[2] * u[2];
    lam = (nr0[0] * (pt0[0] - xt[0]) + nr0[1] * (pt0[1] - xt[1]) +
           nr0[2] * (pt0[2] - xt[2])) * rcp64(nu);
  }
  // the polyline at vn: surface crossing, pads, residual, sums
  const double vvn = vn[0] * vn[0] + vn[1] * vn[1];
  const double vm = sqrt64(vvn);
  bool fin = lay_ok && isfinite(c1n) && isfinite(a) && isfinite(rN);
  bool sides = N > 1 ? (Ls.dpos != 0 && dN > 0.0) : true;
  bool evan = N > 1 && vm > Ls.min_n_evan;
  double s[3];
  for (int c = 0; c < 3; ++c) s[c] = xt[c] + lam * u[c];
  fin = fin && isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]);
  for (int i = 0; i < kp; ++i) {
    const double zp = rc.desc[i * SS_IDESC_COUNT + SS_IDESC_MP + 2];
    sides = sides && ((p[2] - zp) * (s[2] - zp) < 0.0);
  }
  if (kp > 0) {
    const double zp = rc.desc[SS_IDESC_MP + 2];
    const double f = (p[2] - zp) * rcp64(p[2] - s[2]);
    for (int c = 0; c < 3; ++c) ps.x_first[c] = p[c] + f * (s[c] - p[c]);
    ps.x_first[2] = zp;
  } else {
    for (int c = 0; c < 3; ++c) ps.x_first[c] = s[c];
  }
  const double ds[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
  const double sa2 = ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2];
  const double isa = rsq64(sa2 > kTiny * kTiny ? sa2 : kTiny * kTiny);
  const double sa = sa2 * isa;
  const double din[3] = {ds[0] * isa, ds[1] * isa, ds[2] * isa};
  const double in1 = rcp64(nl[1]);
  const double dout[3] = {vn[0] * in1, vn[1] * in1, -c1n * in1};
  const double ct1 = nr0[0] * din[0] + nr0[1] * din[1] + nr0[2] * din[2];
  const double ct2 = nr0[0] * dout[0] + nr0[1] * dout[1] + nr0[2] * dout[2];
  const double t1[3] = {din[0] - ct1 * nr0[0], din[1] - ct1 * nr0[1], din[2] - ct1 * nr0[2]};
  const double st1 = sqrt64(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  const double dG[3] = {nl[0] * din[0] - nl[1] * dout[0], nl[0] * din[1] - nl[1] * dout[1],
                        nl[0] * din[2] - nl[1] * dout[2]};
  const double nd = nr0[0] * dG[0] + nr0[1] * dG[1] + nr0[2] * dG[2];
  const double rt[3] = {dG[0] - nd * nr0[0], dG[1] - nd * nr0[1], dG[2] - nd * nr0[2]};
  const double resid = sqrt64(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
  fin = fin && isfinite(resid) && isfinite(sa);
  sides = sides &&
          ((nr0[0] * (p[0] - s[0]) + nr0[1] * (p[1] - s[1]) + nr0[2] * (p[2] - s[2])) *
               (-lam * nu) < 0.0);
  evan = evan || (nl[0] * st1 > nl[1] * (1.0 + 64.0 * kEps64));
  // surface (medium 0 -> 1, general form) in float32
  {
    const float ci_raw = (float)fabs(ct1), ct_raw = (float)fabs(ct2);
    const float ci = fmaxf(ci_raw, (float)kCMin);
    const float e1 = (float)rc.legs[(rc.n_crossed + 1) + kp], e2 = (float)rc.legs[(rc.n_crossed + 1) + kp + 1];
    const float n0f = (float)nl[0], n1f = (float)nl[1];
    const float sin2 = (e1 * f_rcp(e2)) * (1.0f - ci * ci);
    const float cc2 = f_sqrt(fmaxf(1.0f - sin2, 0.0f));
    const float gm = (n0f * ci - n1f * cc2) * f_rcp(n0f * ci + n1f * cc2);
    tau2 *= 1.0f - gm * gm;
    if (rough) {
      const float sk = Ls.sigf[0] * (k0f * (n0f * ci_raw - n1f * ct_raw));
      tau2 *= f_exp(-2.0f * (sk * sk));
    }
    const float csurf = ci;
    loss += (float)sa * Ls.attf[0];
    perp += (float)sa * f_rcp(n0f);
    par += (float)sa * f_rcp(n0f * csurf * csurf);
    ps.c_first = csurf;
  }
  double opl = nl[0] * sa;
  if (N > 1) {
    // layer 1: the climb |lam| n_1; then the layers' optical path
    const double alam = fabs(lam);
    opl += Ls.n2[1] * alam + o;
    const float len = (float)(alam * nl[1]);
    const float ci = fmaxf(c1f * f_rcp((float)nl[1]), (float)kCMin);
    loss += len * Ls.attf[1];
    perp += (float)alam;
    par += len * f_rcp((float)nl[1] * ci * ci);
    ps.cur[0] = q[0] - dN * vn[0] * rN;
    ps.cur[1] = q[1] - dN * vn[1] * rN;
    ps.cur[2] = Ls.zlast;
    ps.c_last = fmaxf(c_lastf * f_rcp((float)nl[N]), (float)kCMin);
  } else {
    for (int c = 0; c < 3; ++c) ps.cur[c] = s[c];
    ps.c_last = fmax(fabs(ct2), kCMin);
  }
  ps.valid = (resid <= 1e-9) && fin && sides && !evan;
  ps.opl = opl;
  ps.loss_db = loss;
  ps.sum_par = par;
  ps.sum_perp = perp;
  ps.tau2 = tau2;
  return stuck ? !ps.valid : (ps.valid && resid <= 1e-12);
}

// Top-level refracted pair: crossing solver (NMAX < 0: the flat-stack joint
// solve; otherwise the sequential chain, or the joint solve when rc.joint
// with a stack at most NMAX deep) then radiometry.
template <int NMAX, typename FT>
SS_HD void joint_sums(const FT& f, const TraceRec& t, const RefrConst& rc,
                      PathSums& ps) {
  if constexpr (NMAX < 0) {
    flat_path(t.p, f.c, rc, ps);
  } else {
    joint_path<NMAX>(t.p, f.c, rc, rc.k_pad, rc.n_newton, rc.n_backtrack, ps.cur, ps.valid,
                     ps.opl, ps.loss_db, ps.sum_par, ps.sum_perp, ps.tau2, ps.c_first,
                     ps.c_last, ps.x_first);
  }
}

template <int NMAX, int RM = -1, int PAT = -1>
SS_HD void refracted_pair_fp64(const RefrFacet& f, const TraceRec& t, const RefrConst& rc,
                               PairOut& o) {
  PathSums ps;
  if constexpr (NMAX < 0)
    joint_sums<-1>(f, t, rc, ps);
  else if (NMAX > 0 && rc.joint)
    joint_sums<(NMAX > 0 ? NMAX : 1)>(f, t, rc, ps);
  else
    seq_path_fp64(f, t, rc, ps);
  radiometry_fp64<RM, PAT>(f, t, rc, ps, o);
}

// Deepest stack the mixed tier's joint solve takes; deeper stacks keep the
// float64 joint solve (their per-lane arrays live in scratch memory, where
// the float32 search's extra arrays cost more than the float64 arithmetic
// it saves).
constexpr int kMixedJointMax = 8;

// Mixed tier: crossings by the float32-searched joint solve (stacks up to
// kMixedJointMax deep), the float64 joint solve (deeper stacks), the
// flat-stack form (NMAX < 0, layer constants read from ``fl``, staged in LDS
// by the caller) or the sequential chain; radiometry in the mixed tier.
template <int NMAX, int RM = -1, int PAT = -1, typename FT = RefrFacet, bool STAGED = false>
SS_HD bool refracted_pair_mixed(const FT& f, const TraceRec& t, const RefrConst& rc,
                                PairOut& o, const FlatLds* fl = nullptr) {
  PathSums ps;
  if constexpr (NMAX < 0) {
    if (!flat_path_mixed(t.p, f.c, rc, *fl, ps)) return false;
  } else if constexpr (NMAX > kMixedJointMax) {
    if (rc.joint)
      joint_sums<NMAX>(f, t, rc, ps);
    else if (!seq_path_mixed(f, t, rc, ps))
      return false;
  } else {
    if (NMAX > 0 && rc.joint) {
      if (!joint_path_mixed<(NMAX > 0 ? NMAX : 1)>(t.p, f.c, rc, ps)) return false;
    } else if (!seq_path_mixed(f, t, rc, ps)) {
      return false;
    }
  }
  return radiometry_mixed<RM, PAT, FT, STAGED>(f, t, rc, ps, o);
}

}  // namespace ss

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

SS_HD void plane_basis(const double n[3], double b[2][3]) {
  double a[3] = {0.0, 0.0, 0.0};
  if (fabs(n[2]) < 0.9)
    a[2] = 1.0;
  else
    a[0] = 1.0;
  double u[3] = {a[1] * n[2] - a[2] * n[1], a[2] * n[0] - a[0] * n[2],
                 a[0] * n[1] - a[1] * n[0]};
  double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  un = un > kTiny ? un : kTiny;
  for (int k = 0; k < 3; ++k) u[k] /= un;
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

// merit |F|^2 of offsets ab (+ lam * delta) and optionally F itself.
template <int NMAX>
SS_HD double joint_geom(JointWork<NMAX>& w, int N, const double ps[3], const double qs[3],
                        const double* nseg, double lam, bool with_delta, bool keep_f,
                        double s_out[NMAX + 1], double dh_out[NMAX + 1][3]) {
  double prev[3] = {ps[0], ps[1], ps[2]};
  double gprev[3] = {0, 0, 0};
  double merit = 0.0;
  for (int k = 0; k <= N; ++k) {
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
  for (int i = 0; i < N; ++i) {
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
  for (int i = 0; i < N; ++i) {
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
    // blocks
    double wprev[6];
    mat3_sym_outer(dhk[0], nseg[0] / sk[0], wprev);
    for (int i = 0; i < N; ++i) {
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
    for (int i = 0; i < N; ++i) {
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
    for (int i = N - 1; i >= 0; --i) {
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
    for (int i = 0; i < N; ++i) {
      w.ab[i][0] += sel * w.delta[i][0];
      w.ab[i][1] += sel * w.delta[i][1];
    }
  }
  // final crossings (unscaled) and the honest polyline geometry
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 3; ++c)
      x[i][c] = (w.o[i][c] + (w.B[i][0][c] * w.ab[i][0] + w.B[i][1][c] * w.ab[i][1])) * scale;
  double prev[3] = {p[0], p[1], p[2]};
  double din[3], dout[3];
  double resid = 0.0;
  bool sides = true, evan = false, fin = true;
  double sdir[NMAX + 1][3];
  for (int k = 0; k <= N; ++k) {
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
  for (int i = 0; i < N; ++i) {
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
  for (int i = 0; i < N; ++i) {
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
  for (int i = 0; i < N; ++i) {
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

// ------------------------------------------------------------ flat stack
// Joint solve specialised to a stack whose interfaces below the surface are
// horizontal planes.  The joint anchoring of such a plane cannot depend on
// where the chain crosses it, so only the surface is anchored: pass 1 of the
// joint solve runs the chain through the no-op pads (each a straight step
// toward q, so the last pad crossing lies on the segment p -> q) and then
// Snell against the surface mean plane (10 iterations) and the cell lookup,
// exactly as joint_path does.  The stationary path then has two unknowns,
// the surface crossing s on the anchored plane.  Below s the horizontal
// slowness v is conserved; it solves H(v; s) = sum_k d_k v / c_k - (q_h -
// s_h) = 0 (c_k = sqrt(n_k^2 - |v|^2), d_k the layer thicknesses under s)
// along u = (q_h - s_h)/R in t = tan(theta) of the slowest layer, where each
// term of g(t) is increasing and concave, so Newton from below converges
// monotonically.  Snell at the surface, F(s) = B^T (n_0 (s-p)/|s-p| - k_1)
// = 0, is solved by damped Newton with the implicit derivative dv/ds =
// -M^-1 [I | v/c_1], M = dH/dv.  The pads lie on the straight air leg and
// change none of the sums (index-matched, booked with the surface cosine),
// so the air leg is booked as one segment.  Validity and path sums follow
// joint_crossings / joint_path on the polyline.

// z of the horizontal interface with absolute index i: its first table cell.
SS_HD double flat_z(const RefrConst& rc, int i) {
  return rc.tables[(long)rc.desc[i * SS_IDESC_COUNT + SS_IDESC_TABOFF] * 6 + 2];
}

struct FlatStack {
  const RefrConst* rc;
  const double* n;  // media indices from the air leg down (N + 1)
  int kp, N;        // pads, real interfaces (surface first)
  double nm;        // slowest medium below the surface
  // thickness of real layer k (1..N) under a surface crossing at height sz
  SS_HD double d(int k, double sz, double qz) const {
    const double top = k == 1 ? sz : flat_z(*rc, kp + k - 1);
    const double bot = k == N ? qz : flat_z(*rc, kp + k);
    return top - bot;
  }
};

// Horizontal slowness v below s toward q.  t: tangent in the slowest layer
// (in: start, clamped to >= 0; out: solution).  Also c_1 and M^-1.
SS_HD void flat_slowness(const FlatStack& fs, const double s[3], const double q[3],
                         int iters, double& t, double v[2], double& c1,
                         double Minv[2][2]) {
  const double Dx = q[0] - s[0], Dy = q[1] - s[1];
  const double R = sqrt(Dx * Dx + Dy * Dy);
  const double nm = fs.nm;
  t = t > 0.0 ? t : 0.0;
  for (int it = 0; it < iters; ++it) {
    double g = -R, gp = 0.0;
    for (int k = 1; k <= fs.N; ++k) {
      const double dk = fs.d(k, s[2], q[2]), nk = fs.n[k];
      const double den = nk * nk * (1.0 + t * t) - nm * nm * t * t;
      const double sq = sqrt(den);
      g += dk * nm * t / sq;
      gp += dk * nm * nk * nk / (den * sq);
    }
    const double tn = t - g / gp;
    t = tn > 0.0 ? tn : 0.0;
  }
  const double pr = nm * t / sqrt(1.0 + t * t);
  const double ux = R > 0.0 ? Dx / R : 0.0, uy = R > 0.0 ? Dy / R : 0.0;
  v[0] = pr * ux;
  v[1] = pr * uy;
  double a = 0.0, b = 0.0;  // sum d/c, sum d/c^3
  for (int k = 1; k <= fs.N; ++k) {
    const double dk = fs.d(k, s[2], q[2]), nk = fs.n[k];
    const double ck = sqrt(nk * nk - pr * pr);
    a += dk / ck;
    b += dk / (ck * ck * ck);
    if (k == 1) c1 = ck;
  }
  const double m00 = a + b * v[0] * v[0], m01 = b * v[0] * v[1], m11 = a + b * v[1] * v[1];
  const double det = m00 * m11 - m01 * m01;
  Minv[0][0] = m11 / det;
  Minv[1][1] = m00 / det;
  Minv[0][1] = Minv[1][0] = -m01 / det;
}

// Newton state at one surface crossing.
struct FlatState {
  double s[3], t, v[2], c1, Minv[2][2], dh[3], L, F[2];
};

SS_HD void flat_eval(const FlatStack& fs, const double p[3], const double q[3],
                     const double B[2][3], int iters, FlatState& st) {
  flat_slowness(fs, st.s, q, iters, st.t, st.v, st.c1, st.Minv);
  const double n0 = fs.n[0];
  const double ds[3] = {st.s[0] - p[0], st.s[1] - p[1], st.s[2] - p[2]};
  double L = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  L = L > kTiny ? L : kTiny;
  st.L = L;
  for (int c = 0; c < 3; ++c) st.dh[c] = ds[c] / L;
  const double g[3] = {n0 * st.dh[0] - st.v[0], n0 * st.dh[1] - st.v[1],
                       n0 * st.dh[2] + st.c1};
  st.F[0] = B[0][0] * g[0] + B[0][1] * g[1] + B[0][2] * g[2];
  st.F[1] = B[1][0] * g[0] + B[1][1] * g[1] + B[1][2] * g[2];
}

SS_HD void flat_path(const double p[3], const double q[3], const RefrConst& rc, int n_newton,
                     int n_backtrack, double cur[3], bool& valid, double& opl,
                     double& loss_db, double& sum_par, double& sum_perp, double& tau2,
                     double& c_first, double& c_last, double x_first[3]) {
  const int NA = rc.n_crossed, kp = rc.k_pad, N = NA - kp;
  FlatStack fs;
  fs.rc = &rc;
  fs.n = rc.legs + kp;
  fs.kp = kp;
  fs.N = N;
  const double* nl = fs.n;
  const double* epsl = rc.legs + (NA + 1) + kp;
  const double* attl = rc.legs + 2 * (NA + 1) + kp;
  double nm = nl[1];
  for (int k = 2; k <= N; ++k) nm = nl[k] < nm ? nl[k] : nm;
  fs.nm = nm;
  // pass 1: the pads step straight toward q, then the surface mean plane
  double c0[3] = {p[0], p[1], p[2]};
  if (kp > 0) {
    const double zp = rc.desc[(kp - 1) * SS_IDESC_COUNT + SS_IDESC_MP + 2];
    const double f = (p[2] - zp) / (p[2] - q[2]);
    for (int c = 0; c < 3; ++c) c0[c] = p[c] + f * (q[c] - p[c]);
    c0[2] = zp;
  }
  const double* d0 = rc.desc + kp * SS_IDESC_COUNT;
  double pt0[3], nr0[3];
  FlatState st;
  {
    double mp[3] = {d0[SS_IDESC_MP], d0[SS_IDESC_MP + 1], d0[SS_IDESC_MP + 2]};
    double mn[3] = {d0[SS_IDESC_MN], d0[SS_IDESC_MN + 1], d0[SS_IDESC_MN + 2]};
    Snell64 sn;
    snell_fp64(c0, q, mp, mn, nl[0], nl[1], kMeanIters, sn);
    const double* cf = d0 + SS_IDESC_COEF;
    double rc0 = (sn.x[0] * cf[0] + sn.x[1] * cf[2]) + cf[4];
    double rc1 = (sn.x[0] * cf[1] + sn.x[1] * cf[3]) + cf[5];
    int row = clamp_round(rc0, (int)d0[SS_IDESC_NY] - 1);
    int col = clamp_round(rc1, (int)d0[SS_IDESC_NX] - 1);
    const double* g = rc.tables + ((long)d0[SS_IDESC_TABOFF] + (long)row * (int)d0[SS_IDESC_NX] +
                                   col) * 6;
    double nn = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    for (int c = 0; c < 3; ++c) {
      pt0[c] = g[c];
      nr0[c] = g[3 + c] / nn;
    }
    double h = (sn.x[0] - pt0[0]) * nr0[0] + (sn.x[1] - pt0[1]) * nr0[1] +
               (sn.x[2] - pt0[2]) * nr0[2];
    for (int c = 0; c < 3; ++c) st.s[c] = sn.x[c] - h * nr0[c];
  }
  double B[2][3];
  plane_basis(nr0, B);
  const double z_top = N > 1 ? flat_z(rc, kp + 1) : q[2];
  st.t = 0.0;
  flat_eval(fs, p, q, B, 12, st);
  // damped Newton on the surface crossing: a step is halved until it lowers
  // |F|^2 and keeps s above the first horizontal plane; a step that never
  // does ends the solve
  for (int it = 0; it < n_newton; ++it) {
    const double m0 = st.F[0] * st.F[0] + st.F[1] * st.F[1];
    if (!(m0 > 1e-30)) break;
    double dvds[2][3];
    for (int r = 0; r < 2; ++r) {
      dvds[r][0] = -st.Minv[r][0];
      dvds[r][1] = -st.Minv[r][1];
      dvds[r][2] = -(st.Minv[r][0] * st.v[0] + st.Minv[r][1] * st.v[1]) / st.c1;
    }
    double G[3][3];
    const double n0 = nl[0];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double pm = ((r == c ? 1.0 : 0.0) - st.dh[r] * st.dh[c]) / st.L;
        const double dk = r < 2 ? dvds[r][c]
                                : (st.v[0] * dvds[0][c] + st.v[1] * dvds[1][c]) / st.c1;
        G[r][c] = n0 * pm - dk;
      }
    double J[2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        double acc = 0.0;
        for (int r = 0; r < 3; ++r) {
          const double gb = G[r][0] * B[b][0] + G[r][1] * B[b][1] + G[r][2] * B[b][2];
          acc += B[a][r] * gb;
        }
        J[a][b] = acc;
      }
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double st0 = (J[1][1] * st.F[0] - J[0][1] * st.F[1]) / det;
    const double st1 = (J[0][0] * st.F[1] - J[1][0] * st.F[0]) / det;
    double lam = 1.0;
    bool taken = false;
    for (int bt = 0; bt < n_backtrack && !taken; ++bt, lam *= 0.5) {
      FlatState tr = st;
      for (int c = 0; c < 3; ++c) tr.s[c] = st.s[c] - lam * (B[0][c] * st0 + B[1][c] * st1);
      flat_eval(fs, p, q, B, 6, tr);
      const double mt = tr.F[0] * tr.F[0] + tr.F[1] * tr.F[1];
      if (isfinite(mt) && mt < m0 && tr.s[2] > z_top) {
        st = tr;
        taken = true;
      }
    }
    if (!taken) break;
  }
  double Minv[2][2];
  flat_slowness(fs, st.s, q, 4, st.t, st.v, st.c1, Minv);
  // polyline, streamed top-down: validity and sums as joint_path
  const double* s = st.s;
  const double* v = st.v;
  double prev[3] = {p[0], p[1], p[2]};
  double xi[3] = {s[0], s[1], s[2]};
  double resid = 0.0, c_surf = 1.0, ct2_last = 1.0;
  bool sides = true, evan = false, fin = true;
  tau2 = 1.0;
  opl = loss_db = sum_par = sum_perp = 0.0;
  for (int i = 0; i < kp; ++i) {
    const double zp = rc.desc[i * SS_IDESC_COUNT + SS_IDESC_MP + 2];
    sides = sides && ((p[2] - zp) * (s[2] - zp) < 0.0);
  }
  if (kp > 0) {
    const double zp = rc.desc[SS_IDESC_MP + 2];
    const double f = (p[2] - zp) / (p[2] - s[2]);
    for (int c = 0; c < 3; ++c) x_first[c] = p[c] + f * (s[c] - p[c]);
    x_first[2] = zp;
  } else {
    for (int c = 0; c < 3; ++c) x_first[c] = s[c];
  }
  const double pr2 = v[0] * v[0] + v[1] * v[1];
  for (int i = 0; i < N; ++i) {
    double nxt[3];
    if (i + 1 < N) {
      const double dk = fs.d(i + 1, s[2], q[2]), nk = nl[i + 1];
      const double ck = sqrt(nk * nk - pr2);
      nxt[0] = xi[0] + dk * v[0] / ck;
      nxt[1] = xi[1] + dk * v[1] / ck;
      nxt[2] = flat_z(rc, kp + i + 1);
    } else {
      nxt[0] = q[0];
      nxt[1] = q[1];
      nxt[2] = q[2];
    }
    const double n[3] = {i == 0 ? nr0[0] : 0.0, i == 0 ? nr0[1] : 0.0, i == 0 ? nr0[2] : 1.0};
    const double a[3] = {xi[0] - prev[0], xi[1] - prev[1], xi[2] - prev[2]};
    const double b[3] = {nxt[0] - xi[0], nxt[1] - xi[1], nxt[2] - xi[2]};
    const double sa = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double sb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double sam = sa > kTiny ? sa : kTiny, sbm = sb > kTiny ? sb : kTiny;
    const double din[3] = {a[0] / sam, a[1] / sam, a[2] / sam};
    const double dout[3] = {b[0] / sbm, b[1] / sbm, b[2] / sbm};
    const double ct1 = n[0] * din[0] + n[1] * din[1] + n[2] * din[2];
    const double ct2 = n[0] * dout[0] + n[1] * dout[1] + n[2] * dout[2];
    const double t1[3] = {din[0] - ct1 * n[0], din[1] - ct1 * n[1], din[2] - ct1 * n[2]};
    const double t2[3] = {dout[0] - ct2 * n[0], dout[1] - ct2 * n[1], dout[2] - ct2 * n[2]};
    const double st1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    const double st2 = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    const double th1 = atan2(st1, fabs(ct1)), th2 = atan2(st2, fabs(ct2));
    const double dG[3] = {nl[i] * din[0] - nl[i + 1] * dout[0],
                          nl[i] * din[1] - nl[i + 1] * dout[1],
                          nl[i] * din[2] - nl[i + 1] * dout[2]};
    const double nd = n[0] * dG[0] + n[1] * dG[1] + n[2] * dG[2];
    const double rt[3] = {dG[0] - nd * n[0], dG[1] - nd * n[1], dG[2] - nd * n[2]};
    const double rn = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    resid = rn > resid ? rn : resid;
    fin = fin && isfinite(rn) && isfinite(sa) && isfinite(xi[0]) && isfinite(xi[1]) &&
          isfinite(xi[2]);
    const double hp = -(n[0] * a[0] + n[1] * a[1] + n[2] * a[2]);
    const double hn = n[0] * b[0] + n[1] * b[1] + n[2] * b[2];
    sides = sides && (hp * hn < 0.0);
    evan = evan || (nl[i] * st1 > nl[i + 1] * (1.0 + 64.0 * kEps64));
    const double ci = fmax(cos(th1), kCMin);
    if (i == 0) c_surf = ci;
    const double e1 = epsl[i], e2 = epsl[i + 1];
    const double n1 = sqrt(e1), n2 = sqrt(e2);
    const double sin2 = (e1 / e2) * (1.0 - ci * ci);
    const double cc2 = sqrt(fmax(1.0 - sin2, 0.0));
    const double gm = (n1 * ci - n2 * cc2) / (n1 * ci + n2 * cc2);
    tau2 *= (1.0 - gm * gm);
    if (rc.flags & SS_CROSS_ROUGH) {
      const double kt = rc.k0 * (nl[i] * cos(th1) - nl[i + 1] * cos(th2));
      const double sk = rc.desc[(kp + i) * SS_IDESC_COUNT + SS_IDESC_SIGC] * kt;
      tau2 *= exp(-2.0 * (sk * sk));
    }
    opl += nl[i] * sa;
    loss_db += sa * (attl[i] / 1000.0);
    sum_perp += sa / nl[i];
    sum_par += sa / (nl[i] * ci * ci);
    ct2_last = cos(th2);
    for (int c = 0; c < 3; ++c) prev[c] = xi[c];
    if (i + 1 < N)
      for (int c = 0; c < 3; ++c) xi[c] = nxt[c];
  }
  valid = (resid <= 1e-9) && fin && sides && !evan;
  c_first = c_surf;
  c_last = fmax(ct2_last, kCMin);
  for (int c = 0; c < 3; ++c) cur[c] = xi[c];
}

// Top-level refracted pair: crossing solver (NMAX < 0: the flat-stack joint
// solve; otherwise the sequential chain, or the joint solve when rc.joint
// with a stack at most NMAX deep) then radiometry.
template <int NMAX, typename FT>
SS_HD void joint_sums(const FT& f, const TraceRec& t, const RefrConst& rc,
                      PathSums& ps) {
  if constexpr (NMAX < 0) {
    flat_path(t.p, f.c, rc, rc.n_newton, rc.n_backtrack, ps.cur, ps.valid, ps.opl,
              ps.loss_db, ps.sum_par, ps.sum_perp, ps.tau2, ps.c_first, ps.c_last, ps.x_first);
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

// Mixed tier: the joint solve itself stays float64 (its damped Newton needs
// the float64 merit comparisons); radiometry runs in the mixed tier.
template <int NMAX, int RM = -1, int PAT = -1, typename FT = RefrFacet>
SS_HD bool refracted_pair_mixed(const FT& f, const TraceRec& t, const RefrConst& rc,
                                PairOut& o) {
  PathSums ps;
  if constexpr (NMAX < 0) {
    joint_sums<-1>(f, t, rc, ps);
  } else {
    if (NMAX > 0 && rc.joint)
      joint_sums<(NMAX > 0 ? NMAX : 1)>(f, t, rc, ps);
    else if (!seq_path_mixed(f, t, rc, ps))
      return false;
  }
  return radiometry_mixed<RM, PAT, FT>(f, t, rc, ps, o);
}

}  // namespace ss

// CPU build of the device pair math (radar_sounder_simulator_amd/csrc/
// ss_math.h, ss_refract.h) for tests: the same functions the gfx950 kernels
// call, compiled with g++ and summed sequentially per trace, so the kernel
// arithmetic is checked against the oracle without a GPU.  Test
// infrastructure only.
#include <cstring>
#include <vector>

#include "../../radar_sounder_simulator_amd/csrc/ss_joint.h"

using namespace ss;

// The incoherent kernel's float32 tier reads a tabulated pattern through
// the lookup it stages in LDS (pattern_lds_value / _index); the harness builds the
// same lookup in host memory.
struct HostPatLookup {
  std::vector<float> ftab;
  std::vector<short> fidx;
};

static void stage_host(PatternArgs& pa, HostPatLookup& st) {
  pa.ftab = nullptr;
  pa.fidx = nullptr;
  pa.fscale = 0.0f;
  if (pa.kind != SS_PAT_TABULATED || pa.n < 2 || pa.n > kPatLds || !pa.table) return;
  st.ftab.assign(3 * pa.n, 0.0f);
  st.fidx.assign(kPatCells, 0);
  for (int i = 0; i < pa.n; ++i) pattern_lds_value(pa.table, pa.n, i, st.ftab.data());
  for (int i = 0; i < kPatCells; ++i) pattern_lds_index(st.ftab.data(), pa.n, i, st.fidx.data());
  pa.ftab = st.ftab.data();
  pa.fidx = st.fidx.data();
  pa.fscale = pattern_cell_scale(pa.table, pa.n);
}

static TraceRec load_trace(const double* tr, long ts, long t) {
  TraceRec r;
  double* d = reinterpret_cast<double*>(&r);
  for (int i = 0; i < SS_TROW_COUNT; ++i) d[i] = tr[(long)i * ts + t];
  return r;
}

static void accumulate(const PairOut& o, long t, int n, int n_seg, bool split,
                       double* field, double* dropped) {
  if (!o.valid) {
    dropped[t] += o.pwr_invalid;
    return;
  }
  for (int p = 0; p < o.nparts; ++p) {
    int b = o.bin[p];
    if (b >= 0 && b < n) {
      long idx = (t * n_seg + (split ? o.right : 0) * n + b) * 2;
      field[idx] += o.re[p];
      field[idx + 1] += o.im[p];
    } else {
      dropped[t] += o.power ? o.re[p] : o.re[p] * o.re[p] + o.im[p] * o.im[p];
    }
  }
}

extern "C" int host_surface(int prec, const ss_surface_params* prm, const double* F, long fs,
                            long nf, const double* tr, long ts, long nt, const double* table,
                            const double* cst, double* field, double* dropped, int* bins) {
  SurfaceConst sc;
  sc.k = prm->k; sc.gamma = prm->gamma; sc.t0 = prm->t0; sc.dt = prm->dt; sc.c = prm->c;
  sc.taper_s = prm->taper_s; sc.sigma = prm->sigma; sc.lc = prm->corr_len;
  sc.n_samples = prm->n_samples; sc.flags = prm->flags; sc.n_terms = prm->n_terms;
  sc.acf = prm->acf;
  sc.pat.kind = prm->pattern; sc.pat.n = prm->pat_n; sc.pat.a = prm->pat_a;
  for (int i = 0; i < 4; ++i) sc.pat.b[i] = prm->pat_b[i];
  sc.pat.table = table; sc.cst = cst;
  HostPatLookup plk;
  stage_host(sc.pat, plk);  // the float32 tier reads a staged tabulated table
  sc.k_over_pi = prm->k / kPi; sc.kfac = prm->k / kTwoPi;
  sc.k_bin = 2.0 / (prm->c * prm->dt); sc.t0_dt = prm->t0 / prm->dt; sc.cyc_per_m = prm->k / kPi;
  bool split = prm->flags & SS_SPLIT_SIDES;
  int n = prm->n_samples, n_seg = (split ? 2 : 1) * n;
  std::memset(field, 0, sizeof(double) * nt * n_seg * 2);
  std::memset(dropped, 0, sizeof(double) * nt);
  for (long f = 0; f < nf; ++f) {
    SurfFacet64 a;
    for (int j = 0; j < 3; ++j) {
      a.c[j] = F[(SS_FROW_CX + j) * fs + f]; a.n[j] = F[(SS_FROW_NX + j) * fs + f];
      a.e1[j] = F[(SS_FROW_E1X + j) * fs + f]; a.e2[j] = F[(SS_FROW_E2X + j) * fs + f];
    }
    a.A = F[SS_FROW_AREA * fs + f];
    a.gam = (prm->flags & SS_GAMMA_FACET) ? F[SS_FROW_GAMMA * fs + f] : 0.0;
    a.ph_re = (prm->flags & SS_ROUGH) ? F[SS_FROW_PH_RE * fs + f] : 0.0;
    a.ph_im = (prm->flags & SS_ROUGH) ? F[SS_FROW_PH_IM * fs + f] : 0.0;
    double l1 = sqrt((a.e1[0] * a.e1[0] + a.e1[1] * a.e1[1]) + a.e1[2] * a.e1[2]);
    double l2 = sqrt((a.e2[0] * a.e2[0] + a.e2[1] * a.e2[1]) + a.e2[2] * a.e2[2]);
    a.l1 = l1 > 0 ? l1 : 1.0; a.l2 = l2 > 0 ? l2 : 1.0;
    SurfFacet32 b;
    for (int j = 0; j < 3; ++j) {
      b.c[j] = a.c[j]; b.n[j] = (float)a.n[j]; b.e1[j] = (float)a.e1[j]; b.e2[j] = (float)a.e2[j];
    }
    b.A = (float)a.A; b.gam = (float)a.gam; b.ph_re = (float)a.ph_re; b.ph_im = (float)a.ph_im;
    b.l1 = (float)a.l1; b.l2 = (float)a.l2;
    for (long t = 0; t < nt; ++t) {
      TraceRec r = load_trace(tr, ts, t);
      PairOut o{};
      if (prec == SS_PREC_FP64) surface_pair_fp64(a, r, sc, o);
      else if (sc.pat.ftab) surface_pair_mixed<-1, -1, true>(b, nullptr, r, sc, o);
      else surface_pair_mixed(b, nullptr, r, sc, o);
      accumulate(o, t, n, n_seg, split, field, dropped);
      if (bins) bins[t * nf + f] = o.raw_bin;
    }
  }
  if (prec == SS_PREC_FP64) {  // fold-back
    for (long t = 0; t < nt; ++t) {
      double th = fmod(2.0 * prm->k * tr[(long)SS_TROW_RREF * ts + t], kTwoPi);
      double c = cos(-th), s = sin(-th);
      for (int i = 0; i < n_seg; ++i) {
        double* v = field + (t * n_seg + i) * 2;
        double a = v[0], b = v[1];
        v[0] = a * c - b * s; v[1] = a * s + b * c;
      }
    }
  }
  return 0;
}

extern "C" int host_refracted(int prec, const ss_refracted_params* prm, const double* F, long fs,
                              long nf, const double* tr, long ts, long nt, const double* desc,
                              const double* tables, const double* legs, const double* table,
                              const double* cst, double* out, double* dropped, int* bins,
                              long* n_fallback) {
  RefrConst rc;
  rc.k0 = prm->k0; rc.gamma = prm->gamma; rc.t0 = prm->t0; rc.dt = prm->dt; rc.c = prm->c;
  rc.taper_s = prm->taper_s; rc.sigma = prm->sigma; rc.lc = prm->corr_len; rc.n_exp = prm->n_exp;
  rc.n_samples = prm->n_samples; rc.flags = prm->flags; rc.n_crossed = prm->n_crossed;
  rc.n_terms = prm->n_terms; rc.acf = prm->acf;
  rc.pat.kind = prm->pattern; rc.pat.n = prm->pat_n; rc.pat.a = prm->pat_a;
  for (int i = 0; i < 4; ++i) rc.pat.b[i] = prm->pat_b[i];
  rc.pat.table = table; rc.cst = cst; rc.desc = desc; rc.tables = tables; rc.legs = legs;
  HostPatLookup plk;
  stage_host(rc.pat, plk);
  // the D_Phi term constants, staged as the kernels stage them in LDS
  std::vector<float> mtab(2 * kTermTab);
  for (int m = 0; m < kTermTab; ++m) term_tab_entry(m, mtab.data());
  rc.mtab = mtab.data();
  rc.k_bin = 2.0 / (prm->c * prm->dt); rc.t0_dt = prm->t0 / prm->dt; rc.cyc_per_m = prm->k0 / kPi;
  rc.fast_opl_tol = (float)(kFastOplFrac * kBinGuard / rc.k_bin);
  rc.joint = prm->refraction == 1; rc.k_pad = prm->k_pad;
  rc.n_newton = prm->joint_newton; rc.n_backtrack = prm->joint_backtrack;
  bool split = prm->flags & SS_SPLIT_SIDES;
  int n = prm->n_samples, n_seg = (split ? 2 : 1) * n;
  std::memset(out, 0, sizeof(double) * nt * n_seg * 2);
  std::memset(dropped, 0, sizeof(double) * nt);
  long fb = 0;
  for (long f = 0; f < nf; ++f) {
    RefrFacet a;
    for (int j = 0; j < 3; ++j) {
      a.c[j] = F[(SS_FROW_CX + j) * fs + f]; a.n[j] = F[(SS_FROW_NX + j) * fs + f];
      a.e1[j] = F[(SS_FROW_E1X + j) * fs + f]; a.e2[j] = F[(SS_FROW_E2X + j) * fs + f];
    }
    a.A = F[SS_FROW_AREA * fs + f];
    a.gam = (prm->flags & SS_GAMMA_FACET) ? F[SS_FROW_GAMMA * fs + f] : 0.0;
    a.ph_re = (prm->flags & SS_ROUGH) ? F[SS_FROW_PH_RE * fs + f] : 0.0;
    a.ph_im = (prm->flags & SS_ROUGH) ? F[SS_FROW_PH_IM * fs + f] : 0.0;
    double l1 = sqrt((a.e1[0] * a.e1[0] + a.e1[1] * a.e1[1]) + a.e1[2] * a.e1[2]);
    double l2 = sqrt((a.e2[0] * a.e2[0] + a.e2[1] * a.e2[1]) + a.e2[2] * a.e2[2]);
    a.l1 = l1 > 0 ? l1 : 1.0; a.l2 = l2 > 0 ? l2 : 1.0;
    a.dif = (prm->flags & SS_DIFFUSE) ? F[SS_FROW_DIF_AMP * fs + f] : 0.0;
    a.dph_re = (prm->flags & SS_DIFFUSE) ? F[SS_FROW_DIF_RE * fs + f] : 0.0;
    a.dph_im = (prm->flags & SS_DIFFUSE) ? F[SS_FROW_DIF_IM * fs + f] : 0.0;
    for (long t = 0; t < nt; ++t) {
      TraceRec r = load_trace(tr, ts, t);
      PairOut o{};
      if (prec == SS_PREC_FP64) refracted_pair_fp64<SS_MAX_JOINT>(a, r, rc, o);
      else if (!(rc.pat.ftab ? refracted_pair_mixed<SS_MAX_JOINT, -1, -1, RefrFacet, true>(a, r, rc, o)
                             : refracted_pair_mixed<SS_MAX_JOINT>(a, r, rc, o))) {
        ++fb;
        refracted_pair_fp64<SS_MAX_JOINT>(a, r, rc, o);
      }
      accumulate(o, t, n, n_seg, split, out, dropped);
      if (bins) bins[t * nf + f] = o.raw_bin;
    }
  }
  if (n_fallback) *n_fallback = fb;
  return 0;
}

// The staged tabulated lookup at given angles, and the float32 acos.
extern "C" void host_tab_lookup(const double* table, int n, const float* th, long m,
                                float* out) {
  PatternArgs pa{};
  pa.kind = SS_PAT_TABULATED;
  pa.n = n;
  pa.table = table;
  HostPatLookup plk;
  stage_host(pa, plk);
  for (long i = 0; i < m; ++i) out[i] = tab_lookup_staged(pa, th[i]);
}

extern "C" void host_f_acos(const float* x, long m, float* out) {
  for (long i = 0; i < m; ++i) out[i] = f_acos(x[i]);
}

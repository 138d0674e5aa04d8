// ss_kernels.hip -- gfx950 kernels and the C ABI of soundersim_hip.h.
//
// Work decomposition (surface and refracted kernels alike): one workgroup of
// 256 lanes = one facet tile (256 consecutive facets of a Morton-ordered
// scene, i.e. a compact 2-D patch) x one trace tile (TT traces).  Each lane
// keeps its facet in registers; the trace tile is staged in LDS and walked
// trace by trace.  For each trace the workgroup
//   1. evaluates its 256 pairs,
//   2. finds the tile's bin range with a wavefront min/max reduction
//      (cross-lane shuffles) and one LDS exchange between the 4 waves,
//   3. accumulates the pairs into a float64 LDS bin window (ds_add_f64)
//      -- a compact patch spans a few to a few tens of bins per trace --
//   4. flushes the occupied window slots to HBM with one float64 atomic add
//      per slot (the flush of trace k overlaps the accumulation of trace
//      k+1 through double-buffered windows, one barrier per trace).
// Contributions beyond the window go straight to a global atomic; out-of-
// window and invalid-path power accumulates per trace in LDS (dropped).
// Block ids are remapped so each XCD gets a contiguous range of
// (trace tile, facet tile) work, keeping a trace tile's output rows in one
// XCD's L2.
//
// Deferred float64 redo (mixed refracted kernels).  A pair whose reduced-
// precision path cannot be resolved to the float64 bin/cell decisions (a
// residual in the ambiguous band, a bin position within the guard of an
// edge) contributes nothing in the hot kernel; each wave records the lanes
// concerned per trace as a 64-bit ballot in LDS, and every work item writes
// one flag (any such lane or not) and, when flagged, its lane masks to the
// caller's workspace -- plain stores, one slot per work item, nothing to
// reset.  A second launch (redo_kernel / the batch redo pass) walks the
// flags and runs the float64 trace loop on the flagged items only, masked
// to the recorded lanes.  The hot kernel never carries the float64 path's
// registers (about 170 vs 300 VGPRs).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "ss_joint.h"

namespace ss {

constexpr int kBlock = 256;
constexpr int kWin = 256;      // LDS slots per side per trace
constexpr int kTTSurf = 32;    // traces per workgroup, surface
constexpr int kTTRefr = 16;    // traces per workgroup, refracted

thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Wave-uniform copy of a value every lane of the workgroup holds (an LDS
// broadcast, a table entry indexed by one), so that loop exits and job
// selection in the persistent kernels compile to scalar branches.
__device__ inline int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline long uniform_l(long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(v & 0xffffffffL));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)((unsigned long)v >> 32));
  return (long)(((unsigned long)hi << 32) | lo);
}

// Wave min / max, called with every lane of the wave active; the result is
// wave-uniform.  On the device: four DPP steps make each 16-lane row uniform
// (quad swaps [1,0,3,2] and [2,3,0,1], half-row mirror, row mirror), then the
// four row values are read into scalars -- no LDS permute round trips.  The
// CPU emulation of the kernels takes the butterfly of lane shuffles.
template <bool MAX>
__device__ inline int wave_ext_i(int v) {
  auto f = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
#if defined(__HIP_DEVICE_COMPILE__)
  v = f(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
  v = f(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
  v = f(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
  v = f(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
  const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return f(f(a, b), f(c, d));
#else
  for (int o = 32; o >= 1; o >>= 1) v = f(v, __shfl_xor(v, o, 64));
  return v;
#endif
}
__device__ inline int wave_min_i(int v) { return wave_ext_i<false>(v); }
__device__ inline int wave_max_i(int v) { return wave_ext_i<true>(v); }

struct KCommon {
  const double* facets;
  long fs;
  long nf;
  const double* traces;
  long ts;
  long nt;
  double* field;
  double* dropped;
  int* pair_bins;
  int n_ftiles, n_ttiles;
  long per_xcd;
  int n_seg;
};

// Work id of this block (XCD-aware remap; -1 when idle).
__device__ inline long work_id(const KCommon& k) {
  long bid = blockIdx.x;
  long w = (bid & 7) * k.per_xcd + (bid >> 3);
  long total = (long)k.n_ftiles * k.n_ttiles;
  return w < total ? w : -1;
}

// Per-workgroup LDS of the trace loop.  Both precision tiers accumulate
// their window in float64: ds_add_f64 measured faster than ds_add_f32 on
// gfx950 for these scatter patterns (surface kernel of the 512 x 512
// surface+bed scene 1.19 -> 1.02 ms), and the sums of <= 256 float32 pair
// contributions per bin carry no float32 rounding.  SIDES: 2 when the
// launch may split sides (both windows), 1 when it cannot.
template <typename ACC, int TT, int SIDES = 2>
struct Shared {
  TraceRec tr[TT];
  ACC win[2][SIDES * kWin][2];
  int wmin[2][4];
  int wmax[2][4];
  double drop[TT];
  unsigned long long redo[TT][4];  // per trace, per wave: lanes left to the redo
};

// Workspace header of the entry points that take a workspace.
// [0..4]: work counters of a batched launch (one per job class), byte
// kWsOverflow: the redo record's overflow word.
constexpr long kWsHeader = 256;
constexpr long kWsOverflow = 64;
constexpr int kJobCodes = 5;  // batched launch classes, see job_code

// Redo record of one launch: ``flags`` holds one byte per work item (1 =
// float64 lanes pending), ``masks`` TT x 4 lane masks per flagged item.
struct Redo {
  unsigned char* flags;
  unsigned long long* masks;
  long cap;  // work items the flags/masks can hold
  unsigned* overflow;  // batched launches: set when an item lies beyond cap
};

__host__ __device__ inline long redo_flag_bytes(long n_items) { return (n_items + 7) & ~7L; }

__host__ __device__ inline Redo redo_view(void* ws, long n_items) {
  Redo r;
  char* b = static_cast<char*>(ws);
  r.flags = reinterpret_cast<unsigned char*>(b + kWsHeader);
  r.masks = reinterpret_cast<unsigned long long*>(b + kWsHeader + redo_flag_bytes(n_items));
  r.cap = n_items;
  r.overflow = reinterpret_cast<unsigned*>(b + kWsOverflow);
  return r;
}

__host__ __device__ inline long redo_bytes(long n_items, int tt) {
  return kWsHeader + redo_flag_bytes(n_items) + 8L * 4 * tt * n_items;
}

// POWER: a real-valued output (n_traces, n_seg) float64 with only the re
// slots of the window in use (incoherent_kernel); else complex interleaved.
template <typename ACC, int TT, bool POWER = false, int SIDES = 2>
__device__ inline void flush_window(Shared<ACC, TT, SIDES>& sh, const KCommon& k, int buf,
                                    long t, int base, int span, int n_samples,
                                    bool split) {
  if (span <= 0) return;
  int w = span < kWin ? span : kWin;
  int nside = (SIDES == 2 && split) ? 2 : 1;
  for (int s = threadIdx.x; s < nside * w; s += kBlock) {
    int side = s / w, slot = s - side * w;
    ACC* v = sh.win[buf][side * kWin + slot];
    if constexpr (POWER) {
      ACC pw = v[0];
      if (pw != (ACC)0) {
        atomicAdd(&k.field[t * k.n_seg + (long)side * n_samples + base + slot], (double)pw);
        v[0] = (ACC)0;
      }
      continue;
    }
    ACC re = v[0], im = v[1];
    if (re != (ACC)0 || im != (ACC)0) {
      long idx = (t * k.n_seg + (long)side * n_samples + base + slot) * 2;
      if (re != (ACC)0) atomicAdd(&k.field[idx], (double)re);
      if (im != (ACC)0) atomicAdd(&k.field[idx + 1], (double)im);
      v[0] = (ACC)0;
      v[1] = (ACC)0;
    }
  }
}

// Trace loop shared by both kernel families.  PAIR(t, o, it) evaluates this
// lane's facet against trace record t (the it-th of the tile) into o; a pair
// left with o.redo set contributes nothing and writes no pair bin.  REDO:
// such pairs are recorded under work item ``item`` of ``rd``.
template <typename ACC, int TT, bool REDO = false, int SIDES = 2, typename PairFn>
__device__ inline void trace_loop(Shared<ACC, TT, SIDES>& sh, const KCommon& k, long w,
                                  bool live, long f, int n_samples, bool split,
                                  PairFn pair, const Redo& rd = Redo{}, long item = 0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long tt_idx = w / k.n_ftiles;
  const long t_base = tt_idx * TT;
  const int nT = (int)((k.nt - t_base) < TT ? (k.nt - t_base) : TT);
  for (int i = tid; i < nT * SS_TROW_COUNT; i += kBlock) {
    int tt = i / SS_TROW_COUNT, r = i - tt * SS_TROW_COUNT;
    reinterpret_cast<double*>(&sh.tr[tt])[r] = k.traces[(long)r * k.ts + t_base + tt];
  }
  for (int i = tid; i < 2 * SIDES * kWin; i += kBlock) {
    ACC* v = sh.win[i / (SIDES * kWin)][i % (SIDES * kWin)];
    v[0] = (ACC)0;
    v[1] = (ACC)0;
  }
  if (tid < TT) sh.drop[tid] = 0.0;
  __syncthreads();
  int prev_base = 0, prev_span = 0;
  for (int it = 0; it < nT; ++it) {
    const int buf = it & 1;
    PairOut o;
    o.nparts = 0;
    o.valid = false;
    o.redo = false;
    o.power = false;
    o.pwr_invalid = 0.0;
    if (live) pair(sh.tr[it], o, it);
    if constexpr (REDO) {
      const unsigned long long m = __ballot(live && o.redo);
      if (lane == 0) sh.redo[it][wave] = m;
    }
    int lmin = INT_MAX, lmax = INT_MIN;
    if (live && o.valid) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {  // fixed trip count keeps o in registers
        if (p >= o.nparts) break;
        int b = o.bin[p];
        if (b >= 0 && b < n_samples) {
          lmin = min(lmin, b);
          lmax = max(lmax, b);
        }
      }
    }
    lmin = wave_min_i(lmin);
    lmax = wave_max_i(lmax);
    if (lane == 0) {
      sh.wmin[buf][wave] = lmin;
      sh.wmax[buf][wave] = lmax;
    }
    __syncthreads();
    if (it > 0)
      flush_window<ACC, TT, false>(sh, k, buf ^ 1, t_base + it - 1, prev_base, prev_span,
                            n_samples, split);
    int base = min(min(sh.wmin[buf][0], sh.wmin[buf][1]),
                   min(sh.wmin[buf][2], sh.wmin[buf][3]));
    int top = max(max(sh.wmax[buf][0], sh.wmax[buf][1]),
                  max(sh.wmax[buf][2], sh.wmax[buf][3]));
    int span = base <= top ? top - base + 1 : 0;
    const long t = t_base + it;
    if (live) {
      if (!o.valid) {
        if (o.pwr_invalid != 0.0) atomicAdd(&sh.drop[it], o.pwr_invalid);
      } else {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          if (p >= o.nparts) break;
          int b = o.bin[p];
          double re = o.re[p], im = o.im[p];
          if (b >= 0 && b < n_samples) {
            int off = b - base;
            int side = (SIDES == 2 && split) ? o.right : 0;
            if (off < kWin) {
              atomicAdd(&sh.win[buf][side * kWin + off][0], (ACC)re);
              atomicAdd(&sh.win[buf][side * kWin + off][1], (ACC)im);
            } else {
              long idx = (t * k.n_seg + (long)side * n_samples + b) * 2;
              if (re != 0.0) atomicAdd(&k.field[idx], re);
              if (im != 0.0) atomicAdd(&k.field[idx + 1], im);
            }
          } else {
            double pw = o.power ? re : re * re + im * im;
            if (pw != 0.0) atomicAdd(&sh.drop[it], pw);
          }
        }
      }
      if (k.pair_bins && !o.redo) k.pair_bins[t * k.nf + f] = o.raw_bin;
    }
    prev_base = base;
    prev_span = span;
  }
  __syncthreads();
  if (nT > 0)
    flush_window<ACC, TT, false>(sh, k, (nT - 1) & 1, t_base + nT - 1, prev_base, prev_span,
                          n_samples, split);
  __syncthreads();
  if (tid < nT && sh.drop[tid] != 0.0) atomicAdd(&k.dropped[t_base + tid], sh.drop[tid]);
  if constexpr (REDO) {
    // every wave's ballot of every trace is visible after the barriers above
    bool any = false;
    for (int i = 0; i < nT; ++i)
      any |= ((sh.redo[i][0] | sh.redo[i][1]) | (sh.redo[i][2] | sh.redo[i][3])) != 0ULL;
    if (item < rd.cap) {  // the caller sizes the record for every item
      if (any && tid < TT * 4)
        rd.masks[item * (TT * 4) + tid] = (tid >> 2) < nT ? sh.redo[tid >> 2][tid & 3] : 0ULL;
      if (tid == 0) rd.flags[item] = any ? 1 : 0;
    }
  }
}

// Kernel-side constants from the ABI parameter structs.
__host__ __device__ inline PatternArgs make_pattern(int kind, int n, double a, const double b[4],
                                                    const double* table) {
  PatternArgs p;
  p.kind = kind;
  p.n = n;
  p.a = a;
  for (int i = 0; i < 4; ++i) p.b[i] = b[i];
  p.table = table;
  p.ftab = nullptr;
  p.fidx = nullptr;
  p.fscale = 0.0f;
  return p;
}

// Stages the float32 lookup of a tabulated pattern of 2 .. kPatLds samples
// (pattern_lds_value, then pattern_lds_index over the staged values) in
// ``ptab`` / ``pidx`` and points ``pa`` at it; visible after the caller's
// next barrier.  Called by every thread of the workgroup (one barrier).
__device__ inline void stage_pattern(PatternArgs& pa, float* ptab, short* pidx) {
  for (int i = threadIdx.x; i < pa.n; i += kBlock) pattern_lds_value(pa.table, pa.n, i, ptab);
  __syncthreads();
  for (int i = threadIdx.x; i < kPatCells; i += kBlock) pattern_lds_index(ptab, pa.n, i, pidx);
  pa.ftab = ptab;
  pa.fidx = pidx;
  pa.fscale = pattern_cell_scale(pa.table, pa.n);
}

// Stages the float32 D_Phi term constants (term_tab_entry, kTermTab entries)
// in ``tab``; visible after the caller's next barrier.
__device__ inline void stage_term_tab(float* tab) {
  for (int m = threadIdx.x; m < kTermTab; m += kBlock) term_tab_entry(m, tab);
}

__host__ __device__ inline SurfaceConst make_surface_const(const ss_surface_params& prm,
                                                           const double* table,
                                                           const double* cst) {
  SurfaceConst sc;
  sc.k = prm.k;
  sc.gamma = prm.gamma;
  sc.t0 = prm.t0;
  sc.dt = prm.dt;
  sc.c = prm.c;
  sc.taper_s = prm.taper_s;
  sc.sigma = prm.sigma;
  sc.lc = prm.corr_len;
  sc.n_samples = prm.n_samples;
  sc.flags = prm.flags;
  sc.n_terms = prm.n_terms;
  sc.acf = prm.acf;
  sc.pat = make_pattern(prm.pattern, prm.pat_n, prm.pat_a, prm.pat_b, table);
  sc.cst = cst;
  sc.k_over_pi = prm.k / kPi;
  sc.kfac = prm.k / kTwoPi;
  sc.k_bin = 2.0 / (prm.c * prm.dt);
  sc.t0_dt = prm.t0 / prm.dt;
  sc.cyc_per_m = prm.k / kPi;
  return sc;
}

__host__ __device__ inline RefrConst make_refr_const(const ss_refracted_params& prm,
                                                     const double* desc, const double* tables,
                                                     const double* legs, const double* table,
                                                     const double* cst) {
  RefrConst rc;
  rc.k0 = prm.k0;
  rc.gamma = prm.gamma;
  rc.t0 = prm.t0;
  rc.dt = prm.dt;
  rc.c = prm.c;
  rc.taper_s = prm.taper_s;
  rc.sigma = prm.sigma;
  rc.lc = prm.corr_len;
  rc.n_exp = prm.n_exp;
  rc.n_samples = prm.n_samples;
  rc.flags = prm.flags;
  rc.n_crossed = prm.n_crossed;
  rc.n_terms = prm.n_terms;
  rc.acf = prm.acf;
  rc.pat = make_pattern(prm.pattern, prm.pat_n, prm.pat_a, prm.pat_b, table);
  rc.cst = cst;
  rc.desc = desc;
  rc.tables = tables;
  rc.legs = legs;
  rc.joint = prm.refraction == 1;
  rc.k_pad = prm.k_pad;
  rc.n_newton = prm.joint_newton;
  rc.n_backtrack = prm.joint_backtrack;
  rc.k_bin = 2.0 / (prm.c * prm.dt);
  rc.t0_dt = prm.t0 / prm.dt;
  rc.cyc_per_m = prm.k0 / kPi;
  rc.fast_opl_tol = (float)(kFastOplFrac * kBinGuard / rc.k_bin);
  return rc;
}

__device__ inline void load_surface_facet(const KCommon& k, long fi, int flags,
                                          SurfFacet64& f64) {
  const double* F = k.facets;
  const long fs = k.fs;
  for (int j = 0; j < 3; ++j) {
    f64.c[j] = F[(SS_FROW_CX + j) * fs + fi];
    f64.n[j] = F[(SS_FROW_NX + j) * fs + fi];
    f64.e1[j] = F[(SS_FROW_E1X + j) * fs + fi];
    f64.e2[j] = F[(SS_FROW_E2X + j) * fs + fi];
  }
  f64.A = F[SS_FROW_AREA * fs + fi];
  f64.gam = (flags & SS_GAMMA_FACET) ? F[SS_FROW_GAMMA * fs + fi] : 0.0;
  f64.ph_re = f64.ph_im = 0.0;
  f64.l1 = f64.l2 = 1.0;
  if (flags & SS_ROUGH) {
    f64.ph_re = F[SS_FROW_PH_RE * fs + fi];
    f64.ph_im = F[SS_FROW_PH_IM * fs + fi];
    double l1 = sqrt((f64.e1[0] * f64.e1[0] + f64.e1[1] * f64.e1[1]) + f64.e1[2] * f64.e1[2]);
    double l2 = sqrt((f64.e2[0] * f64.e2[0] + f64.e2[1] * f64.e2[1]) + f64.e2[2] * f64.e2[2]);
    f64.l1 = l1 > 0.0 ? l1 : 1.0;
    f64.l2 = l2 > 0.0 ? l2 : 1.0;
  }
}

__device__ inline void to_f32(const SurfFacet64& a, SurfFacet32& b) {
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
}

__device__ inline void load_refr_facet(const KCommon& k, long fi, int flags, RefrFacet& fr) {
  const double* F = k.facets;
  const long fs = k.fs;
  for (int j = 0; j < 3; ++j) {
    fr.c[j] = F[(SS_FROW_CX + j) * fs + fi];
    fr.n[j] = F[(SS_FROW_NX + j) * fs + fi];
    fr.e1[j] = F[(SS_FROW_E1X + j) * fs + fi];
    fr.e2[j] = F[(SS_FROW_E2X + j) * fs + fi];
  }
  fr.A = F[SS_FROW_AREA * fs + fi];
  fr.gam = (flags & SS_GAMMA_FACET) ? F[SS_FROW_GAMMA * fs + fi] : 0.0;
  fr.ph_re = fr.ph_im = 0.0;
  fr.l1 = fr.l2 = 1.0;
  if (flags & SS_ROUGH) {
    fr.ph_re = F[SS_FROW_PH_RE * fs + fi];
    fr.ph_im = F[SS_FROW_PH_IM * fs + fi];
    double l1 = sqrt((fr.e1[0] * fr.e1[0] + fr.e1[1] * fr.e1[1]) + fr.e1[2] * fr.e1[2]);
    double l2 = sqrt((fr.e2[0] * fr.e2[0] + fr.e2[1] * fr.e2[1]) + fr.e2[2] * fr.e2[2]);
    fr.l1 = l1 > 0.0 ? l1 : 1.0;
    fr.l2 = l2 > 0.0 ? l2 : 1.0;
  }
  fr.dif = fr.dph_re = fr.dph_im = 0.0;
  if (flags & SS_DIFFUSE) {
    fr.dif = F[SS_FROW_DIF_AMP * fs + fi];
    fr.dph_re = F[SS_FROW_DIF_RE * fs + fi];
    fr.dph_im = F[SS_FROW_DIF_IM * fs + fi];
  }
}

// ----------------------------------------------------------------- surface
// RM: roughness mode (0 none, 1 area-only D_Phi, 2 finite-facet Gaussian,
// -1 any); PAT: antenna kind (0 isotropic, 5 tabulated, -1 any).  The host
// picks the narrowest instantiation so unused features cost no registers.
// FS (mixed tier): 0 every flag at run time; FS_SURF_PLAIN no split sides,
// interpolated bins or per-facet gamma (compile-time), taper and roughness
// at run time.
// The FS_SURF_PLAIN mixed instantiations keep a one-sided bin window (18 KB
// of LDS per workgroup) and are held to 8 waves per SIMD (60-64 VGPRs, no
// spills): the 512 x 512 surface 0.66 -> 0.64 ms against 6 waves.
constexpr int FS_SURF_PLAIN = 1;
constexpr int kSurfPlainRuntimeFlags = SS_TAPER | SS_ROUGH | SS_D_PHI_AREA;
constexpr int kSurfPlainWaves = 8;
template <int PREC, int RM, int PAT, int FS = 0,
          int WAVES = (PREC == SS_PREC_MIXED && FS == FS_SURF_PLAIN) ? kSurfPlainWaves : 1>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void surface_kernel(
    KCommon k, SurfaceConst sc) {
  using ACC = double;  // float64 LDS windows in both tiers (see Shared)
  __shared__ Shared<ACC, kTTSurf, FS == FS_SURF_PLAIN ? 1 : 2> sh;
  const long w = work_id(k);
  if (w < 0) return;
  const long f = (w % k.n_ftiles) * kBlock + threadIdx.x;
  const bool live = f < k.nf;
  SurfFacet64 f64;
  load_surface_facet(k, live ? f : 0, sc.flags, f64);
  const bool split = FS == FS_SURF_PLAIN ? false : (sc.flags & SS_SPLIT_SIDES) != 0;
  if constexpr (PREC == SS_PREC_FP64) {
    trace_loop<ACC, kTTSurf>(sh, k, w, live, f, sc.n_samples, split,
                             [&](const TraceRec& t, PairOut& o, int) {
                               surface_pair_fp64<RM, PAT>(f64, t, sc, o);
                             });
  } else {
    SurfFacet32 f32;
    to_f32(f64, f32);
    // the tabulated instantiation reads its pattern through the lookup
    // staged in LDS (the host picks it for tables of <= kPatLds samples)
    constexpr bool ST = PAT == SS_PAT_TABULATED;
    SurfaceConst scl = sc;
    if constexpr (FS == FS_SURF_PLAIN) scl.flags = sc.flags & kSurfPlainRuntimeFlags;
    if constexpr (ST) {
      __shared__ float ptab[3 * kPatLds];
      __shared__ short pidx[kPatCells];
      stage_pattern(scl.pat, ptab, pidx);
    }
    trace_loop<ACC, kTTSurf>(sh, k, w, live, f, sc.n_samples, split,
                             [&](const TraceRec& t, PairOut& o, int) {
                               surface_pair_mixed<RM, PAT, ST>(f32, nullptr, t, scl, o);
                             });
  }
}

// ------------------------------------------------------------- incoherent
// Power-summing surface kernel (kernels/incoherent.py:31-64).  Its pair is
// ~46 flops, so per-trace work (bin-range reduction, two barriers, window
// flush with one global atomic per occupied bin) and LDS atomics weigh
// more than in the coherent kernels: each lane takes FPL Morton-consecutive
// facets (FPL = 4: a compact 2x2 patch, often one bin), merges equal (bin,
// side) runs in registers before its LDS atomic, and a workgroup covers
// kBlock * FPL facets per trace.  The float32 tier gains from FPL = 4; the
// float64 tier's registers (180 at FPL = 4, 2 waves/SIMD) make FPL = 1
// faster there.  Only the centre, normal and area rows of a facet are read.
template <int PREC>
constexpr int incoherent_fpl() { return PREC == SS_PREC_MIXED ? 4 : 1; }

template <int PREC, int PAT>
__global__ __launch_bounds__(kBlock) void incoherent_kernel(KCommon k, SurfaceConst sc) {
  constexpr int kIncFPL = incoherent_fpl<PREC>();
  using ACC = double;  // float64 LDS windows in both tiers (see Shared)
  constexpr int TT = kTTSurf;
  __shared__ Shared<ACC, TT> sh;
  const long w = work_id(k);
  if (w < 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long fbase = (w % k.n_ftiles) * (long)(kBlock * kIncFPL) + (long)tid * kIncFPL;
  IncohFacet fc[kIncFPL];
  bool live[kIncFPL];
#pragma unroll
  for (int j = 0; j < kIncFPL; ++j) {
    const long f = fbase + j;
    live[j] = f < k.nf;
    const long fi = live[j] ? f : 0;
    for (int q = 0; q < 3; ++q) {
      fc[j].c[q] = k.facets[(SS_FROW_CX + q) * k.fs + fi];
      fc[j].n[q] = k.facets[(SS_FROW_NX + q) * k.fs + fi];
    }
    fc[j].A = k.facets[SS_FROW_AREA * k.fs + fi];
  }
  const bool split = (sc.flags & SS_SPLIT_SIDES) != 0;
  const int n_samples = sc.n_samples;
  const long t_base = (w / k.n_ftiles) * TT;
  const int nT = (int)((k.nt - t_base) < TT ? (k.nt - t_base) : TT);
  for (int i = tid; i < nT * SS_TROW_COUNT; i += kBlock) {
    int tt = i / SS_TROW_COUNT, r = i - tt * SS_TROW_COUNT;
    reinterpret_cast<double*>(&sh.tr[tt])[r] = k.traces[(long)r * k.ts + t_base + tt];
  }
  for (int i = tid; i < 2 * 2 * kWin; i += kBlock) sh.win[i / (2 * kWin)][i % (2 * kWin)][0] = (ACC)0;
  if (tid < TT) sh.drop[tid] = 0.0;
  SurfaceConst scl = sc;
  if constexpr (PREC == SS_PREC_MIXED && PAT == SS_PAT_TABULATED) {
    // float32 gain from the staged lookup (the host picks this
    // instantiation for tables of at most kPatLds samples)
    __shared__ float ptab[3 * kPatLds];
    __shared__ short pidx[kPatCells];
    stage_pattern(scl.pat, ptab, pidx);
  }
  __syncthreads();
  int prev_base = 0, prev_span = 0;
  for (int it = 0; it < nT; ++it) {
    const int buf = it & 1;
    const long t = t_base + it;
    double pw[kIncFPL];
    int bin[kIncFPL], side[kIncFPL];
    int lmin = INT_MAX, lmax = INT_MIN;
#pragma unroll
    for (int j = 0; j < kIncFPL; ++j) {
      PairOut o;
      o.nparts = 0;
      o.power = true;
      if constexpr (PREC == SS_PREC_FP64)
        incoherent_pair_fp64<PAT>(fc[j], sh.tr[it], sc, o);
      else
        incoherent_pair_mixed<PAT>(fc[j], sh.tr[it], scl, o);
      pw[j] = live[j] ? o.re[0] : 0.0;
      bin[j] = o.bin[0];
      side[j] = split ? o.right : 0;
      if (live[j] && bin[j] >= 0 && bin[j] < n_samples) {
        lmin = min(lmin, bin[j]);
        lmax = max(lmax, bin[j]);
      }
      if (k.pair_bins && live[j]) k.pair_bins[t * k.nf + fbase + j] = o.raw_bin;
    }
    lmin = wave_min_i(lmin);
    lmax = wave_max_i(lmax);
    if (lane == 0) {
      sh.wmin[buf][wave] = lmin;
      sh.wmax[buf][wave] = lmax;
    }
    __syncthreads();
    if (it > 0)
      flush_window<ACC, TT, true>(sh, k, buf ^ 1, t - 1, prev_base, prev_span, n_samples,
                                  split);
    const int base = min(min(sh.wmin[buf][0], sh.wmin[buf][1]),
                         min(sh.wmin[buf][2], sh.wmin[buf][3]));
    const int top = max(max(sh.wmax[buf][0], sh.wmax[buf][1]),
                        max(sh.wmax[buf][2], sh.wmax[buf][3]));
    // book: runs of equal (bin, side) merge in registers
    double drop = 0.0, acc = 0.0;
    int cb = 0, cs = 0;
    bool open = false;
#pragma unroll
    for (int j = 0; j <= kIncFPL; ++j) {
      const bool more = j < kIncFPL;
      if (open && (!more || bin[j] != cb || side[j] != cs)) {
        const int off = cb - base;
        if (off < kWin)
          atomicAdd(&sh.win[buf][cs * kWin + off][0], (ACC)acc);
        else
          atomicAdd(&k.field[t * k.n_seg + (long)cs * n_samples + cb], acc);
        open = false;
      }
      if (!more) break;
      if (pw[j] == 0.0) continue;
      if (bin[j] < 0 || bin[j] >= n_samples) {
        drop += pw[j];
      } else if (open) {
        acc += pw[j];
      } else {
        open = true;
        cb = bin[j];
        cs = side[j];
        acc = pw[j];
      }
    }
    if (drop != 0.0) atomicAdd(&sh.drop[it], drop);
    prev_base = base;
    prev_span = base <= top ? top - base + 1 : 0;
  }
  __syncthreads();
  if (nT > 0)
    flush_window<ACC, TT, true>(sh, k, (nT - 1) & 1, t_base + nT - 1, prev_base, prev_span,
                                n_samples, split);
  __syncthreads();
  if (tid < nT && sh.drop[tid] != 0.0) atomicAdd(&k.dropped[t_base + tid], sh.drop[tid]);
}

// Per-trace reference-phase fold-back exp(-j ((2 k r_ref) mod 2 pi)), float64
// surface path (kernels/coherent.py:311-313, 359).
__global__ void fold_kernel(double* field, const double* traces, long ts, long nt,
                            int n_seg, double k) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt * n_seg) return;
  long t = i / n_seg;
  double rref = traces[(long)SS_TROW_RREF * ts + t];
  double th;
  {
#pragma clang fp contract(off)
    th = fmod(2.0 * k * rref, kTwoPi);
  }
  double c = cos(-th), s = sin(-th);
  double a = field[2 * i], b = field[2 * i + 1];
  field[2 * i] = a * c - b * s;
  field[2 * i + 1] = a * s + b * c;
}

// --------------------------------------------------------------- refracted
// NMAX = 0: sequential chain only; NMAX > 0: joint solve for stacks <= NMAX;
// NMAX < 0: flat-stack joint solve.
// The mixed instantiations leave unresolved pairs to redo_kernel (queue rd).
// FS: flag profile.  0: every flag read at run time.  FS_PLAIN (mixed
// tier): the launch is coherent with no split sides, diffuse term,
// per-facet gamma or forced redo -- those flags are compile-time constants,
// and so is the crossing count (one) of a sequential chain, so their tests
// and the chain's crossing loop fold away (the reference likewise compiles
// one variant per combination of static flags); taper, roughness and
// crossing roughness stay run-time.  The host picks it when the launch
// matches.
constexpr int FS_PLAIN = 1;
constexpr int kPlainRuntimeFlags = SS_TAPER | SS_CROSS_ROUGH | SS_ROUGH | SS_D_PHI_AREA;
__host__ __device__ inline bool plain_flags(const ss_refracted_params& p) {
  return (p.flags & SS_COHERENT) && !(p.flags & ~(kPlainRuntimeFlags | SS_COHERENT));
}
__host__ __device__ inline bool plain_profile(const ss_refracted_params& p) {
  return plain_flags(p) && p.n_crossed == 1 && p.refraction == 0;
}

template <int PREC, int NMAX, int RM, int PAT, typename ACC, int FS = 0>
__device__ inline void refracted_body(Shared<ACC, kTTRefr>& sh, const KCommon& k,
                                      const RefrConst& rc, const Redo& rd) {
  const long w = work_id(k);
  if (w < 0) return;
  const long f = (w % k.n_ftiles) * kBlock + threadIdx.x;
  const bool live = f < k.nf;
  const long fi = live ? f : 0;
  const bool split = FS == FS_PLAIN ? false : (rc.flags & SS_SPLIT_SIDES) != 0;
  if constexpr (PREC == SS_PREC_FP64) {
    RefrFacet fr;
    load_refr_facet(k, fi, rc.flags, fr);
    trace_loop<ACC, kTTRefr>(sh, k, w, live, f, rc.n_samples, split,
                             [&](const TraceRec& t, PairOut& o, int) {
                               refracted_pair_fp64<NMAX, RM, PAT>(fr, t, rc, o);
                             });
  } else {
    RefrFacet32 f32;
    {
      RefrFacet fr;
      load_refr_facet(k, fi, rc.flags, fr);
      to_f32(fr, f32);
    }
    constexpr bool ST = PAT == SS_PAT_TABULATED;  // staged pattern lookup
    RefrConst rcl = rc;
    if constexpr (FS == FS_PLAIN) {
      rcl.flags = (rc.flags & kPlainRuntimeFlags) | SS_COHERENT;
      if constexpr (NMAX == 0) rcl.n_crossed = 1;
    }
    if constexpr (ST) {
      __shared__ float ptab[3 * kPatLds];
      __shared__ short pidx[kPatCells];
      stage_pattern(rcl.pat, ptab, pidx);
    }
    if constexpr (RM != 0) {
      __shared__ float mtab[2 * kTermTab];
      stage_term_tab(mtab);
      rcl.mtab = mtab;
    }
    const FlatLds* flp = nullptr;
    if constexpr (NMAX < 0) {
      // flat stack: the layer constants once per workgroup
      __shared__ FlatLds fl;
      stage_flat_layer(rc, fl, threadIdx.x);
      __syncthreads();
      if (threadIdx.x == 0) stage_flat_mins(rc, fl);
      __syncthreads();
      flp = &fl;
    }
    trace_loop<ACC, kTTRefr, true>(sh, k, w, live, f, rc.n_samples, split,
                                   [&](const TraceRec& t, PairOut& o, int) {
                                     if ((rcl.flags & SS_REDO_ALL) ||
                                         !refracted_pair_mixed<NMAX, RM, PAT, RefrFacet32, ST>(
                                             f32, t, rcl, o, flp)) {
                                       o.valid = false;
                                       o.pwr_invalid = 0.0;
                                       o.redo = true;
                                     }
                                   },
                                   rd, w);
  }
}

// The mixed sequential-chain instantiations are held to 3 waves per SIMD
// (left alone the compiler settles just past 168 VGPRs and the kernel runs
// at 2 waves per SIMD), the FS_PLAIN ones to 4 (127 VGPRs with the folded
// flags, no spills; the 512 x 512 bed 2.86 -> 2.51 ms against 3 waves; 5
// waves spill 29 VGPRs to scratch and take 3.08 ms).
template <int PREC, int NMAX, int RM, int PAT, int FS = 0,
          int WAVES = (PREC == SS_PREC_MIXED && NMAX == 0)
                          ? (FS == FS_PLAIN ? 4 : 3)
                          : ((PREC == SS_PREC_MIXED && FS == FS_PLAIN && NMAX > 0 && NMAX <= 3)
                                 ? 2
                                 : 1)>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void refracted_kernel(
    KCommon k, RefrConst rc, Redo rd) {
  using ACC = double;  // float64 LDS windows in both tiers (see Shared)
  __shared__ Shared<ACC, kTTRefr> sh;
  refracted_body<PREC, NMAX, RM, PAT, ACC, FS>(sh, k, rc, rd);
}

// The mixed tier of the flat-stack joint solve (refracted_body with
// NMAX < 0; the float64 tier is refracted_kernel<FP64, -1, ...>).  The
// smooth instantiations are held to 3 waves per SIMD like the sequential
// mixed kernel (the float32 Newton loop, the float64 step and the float64
// pass would otherwise settle just past 168 VGPRs), the smooth FS_PLAIN
// ones to 4 (128 VGPRs and 19 spilled to scratch, yet faster on deep
// stacks: the 24-interface bed of the multilayer scene 13.9 -> 12.5 ms per
// 64 traces), the generic (rough) ones to 2.
template <int PREC, int RM, int PAT, int FS = 0,
          int WAVES = (RM == 0 ? (FS == FS_PLAIN ? 4 : 3) : 2)>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void flat_kernel(
    KCommon k, RefrConst rc, Redo rd) {
  using ACC = double;  // float64 LDS windows in both tiers (see Shared)
  __shared__ Shared<ACC, kTTRefr> sh;
  refracted_body<PREC, -1, RM, PAT, ACC, FS>(sh, k, rc, rd);
}

// Runs ITEM(w) for every flagged work item w < n_items: each block takes
// chunks of kBlock items, one flag per lane, collects the waves' ballots in
// LDS and walks the set bits in order (a uniform loop; ITEM may contain
// barriers).  One pass over the flags costs one byte load per item per lane.
template <typename ItemFn>
__device__ inline void for_flagged_items(const unsigned char* __restrict__ flags, long n_items,
                                         unsigned long long* sh_mask, ItemFn item) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (long base = (long)blockIdx.x * kBlock; base < n_items; base += (long)gridDim.x * kBlock) {
    const long w = base + tid;
    const unsigned long long m = __ballot(w < n_items && flags[w] != 0);
    if (lane == 0) sh_mask[wave] = m;
    __syncthreads();
    for (int wv = 0; wv < kBlock / 64; ++wv) {
      unsigned long long mm = (unsigned long long)uniform_l((long)sh_mask[wv]);
      while (mm != 0ULL) {
        const int bit = __builtin_ctzll(mm);
        mm &= mm - 1ULL;
        item(base + wv * 64 + bit);
      }
    }
    __syncthreads();  // the masks are rewritten by the next chunk
  }
}

// Books one pair's output straight into global memory, exactly as
// trace_loop books it through its LDS window: a valid in-window part to its
// field bin (per side), a valid out-of-window part's power and an invalid
// pair's power to dropped (an incoherent part's re is that power); the raw
// bin to pair_bins.
__device__ inline void book_pair_global(const KCommon& k, const PairOut& o, long t, long f,
                                        int n_samples, bool split) {
  if (!o.valid) {
    if (o.pwr_invalid != 0.0) atomicAdd(&k.dropped[t], o.pwr_invalid);
  } else {
    for (int p = 0; p < 2; ++p) {
      if (p >= o.nparts) break;
      const int b = o.bin[p];
      const double re = o.re[p], im = o.im[p];
      if (b >= 0 && b < n_samples) {
        const int side = split ? o.right : 0;
        const long idx = (t * k.n_seg + (long)side * n_samples + b) * 2;
        if (re != 0.0) atomicAdd(&k.field[idx], re);
        if (im != 0.0) atomicAdd(&k.field[idx + 1], im);
      } else {
        const double pw = o.power ? re : re * re + im * im;
        if (pw != 0.0) atomicAdd(&k.dropped[t], pw);
      }
    }
  }
  if (k.pair_bins) k.pair_bins[t * k.nf + f] = o.raw_bin;
}

// Float64 pass over the pairs the mixed refracted_kernel flagged, compacted:
// each block takes chunks of kBlock work items, counts the set bits of the
// flagged items' lane masks, scans the counts in LDS and hands the chunk's
// pairs out one per lane (item by binary search over the scan, pair by
// rank among the item's set bits, masks in (trace, wave) order).  A pair is
// evaluated with its facet and trace read from global memory and booked
// with global atomics, so the cost follows the number of flagged pairs, not
// of the items holding them.
template <int NMAX, int RM, int PAT>
__global__ __launch_bounds__(kBlock) void redo_kernel(KCommon k, RefrConst rc, Redo rd) {
  constexpr int M = kTTRefr * 4;  // masks per work item
  __shared__ int sh_cnt[kBlock + 1];
  const long n_items = (long)k.n_ftiles * k.n_ttiles;
  const int tid = threadIdx.x;
  const bool split = (rc.flags & SS_SPLIT_SIDES) != 0;
  for (long base = (long)blockIdx.x * kBlock; base < n_items; base += (long)gridDim.x * kBlock) {
    const long w = base + tid;
    int c = 0;
    if (w < n_items && rd.flags[w] != 0)
      for (int i = 0; i < M; ++i) c += __popcll(rd.masks[w * M + i]);
    if (tid == 0) sh_cnt[0] = 0;
    sh_cnt[tid + 1] = c;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {  // inclusive scan of sh_cnt[1..kBlock]
      const int v = tid >= off ? sh_cnt[tid + 1 - off] : 0;
      __syncthreads();
      sh_cnt[tid + 1] += v;
      __syncthreads();
    }
    const int total = uniform_i(sh_cnt[kBlock]);
    for (int r0 = 0; r0 < total; r0 += kBlock) {
      const int g = r0 + tid;
      if (g < total) {
        int lo = 0, hi = kBlock - 1;  // last item whose first pair is <= g
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (sh_cnt[mid] <= g)
            lo = mid;
          else
            hi = mid - 1;
        }
        const long wi = base + lo;
        int rank = g - sh_cnt[lo];
        int mi = 0;
        unsigned long long m = 0ULL;
        for (; mi < M; ++mi) {
          m = rd.masks[wi * M + mi];
          const int pc = __popcll(m);
          if (rank < pc) break;
          rank -= pc;
        }
        for (int j = 0; j < rank; ++j) m &= m - 1ULL;
        const int bit = __builtin_ctzll(m);
        const long f = (wi % k.n_ftiles) * kBlock + (mi & 3) * 64 + bit;
        const long t = (wi / k.n_ftiles) * kTTRefr + (mi >> 2);
        RefrFacet fr;
        load_refr_facet(k, f, rc.flags, fr);
        TraceRec tr;
        for (int r = 0; r < SS_TROW_COUNT; ++r)
          reinterpret_cast<double*>(&tr)[r] = k.traces[(long)r * k.ts + t];
        PairOut o;
        o.nparts = 0;
        o.valid = false;
        o.redo = false;
        o.power = false;
        o.pwr_invalid = 0.0;
        refracted_pair_fp64<NMAX, RM, PAT>(fr, tr, rc, o);
        book_pair_global(k, o, t, f, rc.n_samples, split);
      }
    }
    __syncthreads();  // sh_cnt is rewritten by the next chunk
  }
}

__global__ void reset_workspace_kernel(unsigned long long* hdr) {
  if (threadIdx.x < kJobCodes) hdr[threadIdx.x] = 0ULL;
  if (threadIdx.x == 0) hdr[kWsOverflow / 8] = 0ULL;
}

// ------------------------------------------------------------ batched launch
constexpr int kMaxJobs = 1024;

// A job table in device memory: the job count either from the host (n_host)
// or from device memory (n_dev, the table header of ss_batch_step).
struct BatchArgs {
  const ss_job* jobs;
  const int* n_dev;
  int n_host;
};

__device__ inline int batch_n(const BatchArgs& ba) {
  int n = ba.n_dev ? uniform_i(*ba.n_dev) : ba.n_host;
  return n < 0 ? 0 : (n > kMaxJobs ? kMaxJobs : n);
}

// Zeroes every job's field and dropped power (blockIdx.y strides jobs).
__global__ void batch_zero_kernel(BatchArgs ba) {
  const int n = batch_n(ba);
  for (int j = blockIdx.y; j < n; j += gridDim.y) {
    const ss_job& jb = ba.jobs[j];
    const int ns = jb.kind == 0 ? jb.surface.n_samples : jb.refracted.n_samples;
    const int fl = jb.kind == 0 ? jb.surface.flags : jb.refracted.flags;
    const long nf = jb.n_traces * (long)(((fl & SS_SPLIT_SIDES) ? 2 : 1) * ns) * 2;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf + jb.n_traces;
         i += (long)gridDim.x * blockDim.x) {
      if (i < nf)
        jb.field[i] = 0.0;
      else
        jb.dropped[i - nf] = 0.0;
    }
  }
}

// fold_kernel over every float64 surface job of a batch (blockIdx.y strides
// jobs).
__global__ void batch_fold_kernel(BatchArgs ba) {
  const int n_jobs = batch_n(ba);
  for (int jj = blockIdx.y; jj < n_jobs; jj += gridDim.y) {
    const ss_job& jb = ba.jobs[jj];
    if (jb.kind != 0) continue;
    const int n_seg = ((jb.surface.flags & SS_SPLIT_SIDES) ? 2 : 1) * jb.surface.n_samples;
    const long n = jb.n_traces * (long)n_seg;
    const double k = jb.surface.k;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
      const long t = i / n_seg;
      const double rref = jb.traces[(long)SS_TROW_RREF * jb.trace_stride + t];
      double th;
      {
#pragma clang fp contract(off)
        th = fmod(2.0 * k * rref, kTwoPi);
      }
      const double c = cos(-th), s = sin(-th);
      const double a = jb.field[2 * i], b = jb.field[2 * i + 1];
      jb.field[2 * i] = a * c - b * s;
      jb.field[2 * i + 1] = a * s + b * c;
    }
  }
}

// Work items of one job under trace tile TT.
__host__ __device__ inline long job_work(const ss_job& j, int tt) {
  long nft = (j.n_facets + kBlock - 1) / kBlock;
  long ntt = (j.n_traces + tt - 1) / tt;
  return nft * ntt;
}

// A joint stack (pads included) shallow enough for the register-resident
// joint instantiation.
__host__ __device__ inline bool rc_shallow(const ss_refracted_params& r) {
  return r.n_crossed <= 4;
}

// Launch class of a job: 0 surface; refracted by solver: 1 sequential
// chain, 2 joint solve, 3 joint solve over a flat stack, 4 joint solve over
// a stack of at most kJointTwo interfaces.  Each class runs as its own
// persistent launch (one pair function's registers each: the two-deep
// joint solve needs ~250 VGPRs, the four-deep ~360 VGPRs + AGPRs, so the
// shallow stacks get 2 waves per SIMD where the deeper ones get 1).
constexpr int kJointTwo = 2;
__host__ __device__ inline int job_code(const ss_job& j) {
  if (j.kind == 0) return 0;
  const int r = j.refracted.refraction;
  if (r == 1 && j.refracted.n_crossed <= kJointTwo) return 4;
  return (r >= 0 && r <= 2) ? 1 + r : 1;
}
// Joint stack bound of a refracted launch class (0 sequential, -1 flat).
constexpr int class_nmax(int code) {
  return code == 2 ? SS_MAX_JOINT : (code == 3 ? -1 : (code == 4 ? kJointTwo : 0));
}

// LDS prefix of per-job work counts: begins[i] = first work id of job i.
// CODE >= 0 counts only the jobs of that launch class (the others get no
// items).  The redo record of a mixed launch holds the refracted classes
// one after another: the items of class CODE start at the returned offset
// (the items of the refracted classes below it).
template <int TT, int CODE = -1>
__device__ inline long job_prefix(const ss_job* jobs, int n_jobs, long* begins) {
  const int tid = threadIdx.x;
  __shared__ long s_base;
  for (int i = tid; i < n_jobs; i += kBlock)
    begins[i + 1] = (CODE < 0 || job_code(jobs[i]) == CODE) ? job_work(jobs[i], TT) : 0;
  __syncthreads();
  if (tid == 0) {
    begins[0] = 0;
    long base = 0;
    for (int i = 0; i < n_jobs; ++i) {
      begins[i + 1] += begins[i];
      const int c = job_code(jobs[i]);
      if (CODE > 1 && c >= 1 && c < CODE) base += job_work(jobs[i], TT);
    }
    s_base = base;
  }
  __syncthreads();
  return uniform_l(s_base);
}

// Job owning global work id gw (begins[lo] <= gw < begins[lo + 1]; the
// last such index, so jobs without items are never chosen).
__device__ inline int find_job(const long* begins, int n_jobs, long gw) {
  int lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (begins[mid] <= gw)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <int TT>
__device__ inline KCommon job_common(const ss_job& jb) {
  KCommon k;
  k.facets = jb.facets;
  k.fs = jb.facet_stride;
  k.nf = jb.n_facets;
  k.traces = jb.traces;
  k.ts = jb.trace_stride;
  k.nt = jb.n_traces;
  k.field = jb.field;
  k.dropped = jb.dropped;
  k.pair_bins = nullptr;
  k.n_ftiles = (int)((jb.n_facets + kBlock - 1) / kBlock);
  k.n_ttiles = (int)((jb.n_traces + TT - 1) / TT);
  k.per_xcd = 0;
  const int n = jb.kind == 0 ? jb.surface.n_samples : jb.refracted.n_samples;
  const int fl = jb.kind == 0 ? jb.surface.flags : jb.refracted.flags;
  k.n_seg = ((fl & SS_SPLIT_SIDES) ? 2 : 1) * n;
  return k;
}

// Persistent workgroups drain (job, trace tile, facet tile) items from one
// counter; every workgroup exits once the counter passes the total.  Mixed
// refracted items flag their unresolved pairs under their work id in the
// redo record (each solver class in its own slice of it); an item beyond
// the record's capacity sets the record's overflow word.  RM/PAT: the
// widest roughness mode / antenna kind of the job mix (as in
// surface_kernel), chosen on the host.  KIND: the launch runs the surface
// jobs (0) or the refracted jobs of one solver class (job_code: 1
// sequential chain, 2 joint, 3 flat-stack joint) of the table, so each
// instantiation carries one pair function's registers.  The mixed surface
// and sequential instantiations are held to 3 waves per SIMD (168 VGPRs),
// where the single-launch kernels sit; the job lookup would otherwise push
// them just past it.  The smooth float64 ones are held to 2 (the
// single-launch kernels' occupancy; a few spilled dwords), the rough
// float64 ones keep 1 (2 would spill 100-200 dwords).  The joint classes
// run the generic pair functions: the general joint solve at 1 wave (its
// per-lane Newton state lives in private memory), the flat stack at 2
// (mixed) / 1 (float64), as their single launches.
// Waves per SIMD of a batch_kernel instantiation (see above).  The plain
// mixed ones (FS_PLAIN, feature class 0): surface 6 (78 VGPRs), sequential
// 4 (10 VGPRs spilled; 3 measured slower), flat stack 2 (205 VGPRs; 4
// would spill 64), the joint solve of up to kJointPlain = 3 interfaces 2
// (80 VGPRs spilled, yet faster than 1 wave: the 64-scene mix at 0.25
// scale 523 -> 512 ms; a four-deep bound needs 256 VGPRs + 110 AGPRs and
// runs at 1, 537 ms), the two-deep one 2 (253 VGPRs, no spills).
constexpr int kJointPlain = 3;
constexpr int batch_waves(int prec, int rm, int kind, int fs) {
  return (prec == SS_PREC_MIXED && fs == FS_PLAIN)
             ? (kind == 0 ? 6 : (kind == 1 ? 4 : (kind == 3 ? 2 : 2)))
             : (kind >= 2 ? ((kind == 3 && prec == SS_PREC_MIXED) ? 2 : 1)
                          : (prec == SS_PREC_MIXED ? 3 : (rm == 0 ? 2 : 1)));
}
template <int PREC, int TT, int RM, int PAT, int KIND, int FS = 0,
          int WAVES = batch_waves(PREC, RM, KIND, FS)>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void batch_kernel(BatchArgs ba,
                                                      unsigned long long* counter, Redo rd) {
  using ACC = double;  // float64 LDS windows in both tiers (see Shared)
  __shared__ Shared<ACC, TT> sh;
  __shared__ long begins[kMaxJobs + 1];
  __shared__ long s_w;
  const int tid = threadIdx.x;
  const ss_job* __restrict__ jobs = ba.jobs;
  const int n_jobs = batch_n(ba);
  const long redo_base = job_prefix<TT, KIND>(jobs, n_jobs, begins);
  const long total = uniform_l(begins[n_jobs]);
  int staged = -1;  // job whose flat-stack layers are in LDS (KIND 3, mixed)
  float* mt = nullptr;
  if constexpr (PREC == SS_PREC_MIXED && RM != 0) {
    __shared__ float mtab[2 * kTermTab];  // visible after the loop's first barrier
    stage_term_tab(mtab);
    mt = mtab;
  }
  for (;;) {
    if (tid == 0) s_w = (long)atomicAdd(counter, 1ULL);
    __syncthreads();
    const long gw = uniform_l(s_w);
    __syncthreads();
    if (gw >= total) break;
    const int j = uniform_i(find_job(begins, n_jobs, gw));
    const ss_job& jb = jobs[j];
    const long w = gw - begins[j];
    const KCommon k = job_common<TT>(jb);
    const long f = (w % k.n_ftiles) * kBlock + tid;
    const bool live = f < k.nf;
    if constexpr (KIND == 0) {
      SurfaceConst sc = make_surface_const(jb.surface, jb.pattern_table, jb.constants);
      sc.mtab = mt;
      if constexpr (FS == FS_PLAIN) sc.flags &= kSurfPlainRuntimeFlags;
      const bool split = FS == FS_PLAIN ? false : (sc.flags & SS_SPLIT_SIDES) != 0;
      SurfFacet64 f64;
      load_surface_facet(k, live ? f : 0, sc.flags, f64);
      if constexpr (PREC == SS_PREC_FP64) {
        trace_loop<ACC, TT>(sh, k, w, live, f, sc.n_samples, split,
                            [&](const TraceRec& t, PairOut& o, int) {
                              surface_pair_fp64<RM, PAT>(f64, t, sc, o);
                            });
      } else {
        SurfFacet32 f32;
        to_f32(f64, f32);
        trace_loop<ACC, TT>(sh, k, w, live, f, sc.n_samples, split,
                            [&](const TraceRec& t, PairOut& o, int) {
                              surface_pair_mixed<RM, PAT>(f32, nullptr, t, sc, o);
                            });
      }
    } else {
      // solver of the class: sequential chain, joint solve (stacks up to
      // SS_MAX_JOINT; shallow stacks of at most 4 take the register-resident
      // instantiation: the branch is uniform, one job per work item),
      // flat-stack joint solve
      // (FS_PLAIN: every joint stack is at most kJointPlain deep, the deep
      // solve is not built)
      constexpr int NMAX = (KIND == 2 && FS == FS_PLAIN) ? kJointPlain : class_nmax(KIND);
      const bool shallow = KIND == 2 && (FS == FS_PLAIN || rc_shallow(jb.refracted));
      RefrConst rc = make_refr_const(jb.refracted, jb.iface_desc, jb.tables, jb.legs,
                                     jb.pattern_table, jb.constants);
      rc.mtab = mt;
      if constexpr (FS == FS_PLAIN) {
        rc.flags = (rc.flags & kPlainRuntimeFlags) | SS_COHERENT;
        if constexpr (KIND == 1) rc.n_crossed = 1;
      }
      const bool split = FS == FS_PLAIN ? false : (rc.flags & SS_SPLIT_SIDES) != 0;
      const long fi = live ? f : 0;
      if constexpr (PREC == SS_PREC_FP64) {
        RefrFacet fr;
        load_refr_facet(k, fi, rc.flags, fr);
        trace_loop<ACC, TT>(sh, k, w, live, f, rc.n_samples, split,
                            [&](const TraceRec& t, PairOut& o, int) {
                              if (KIND == 2 && shallow)
                                refracted_pair_fp64<4, RM, PAT>(fr, t, rc, o);
                              else
                                refracted_pair_fp64<NMAX, RM, PAT>(fr, t, rc, o);
                            });
      } else {
        const FlatLds* flp = nullptr;
        if constexpr (KIND == 3) {
          __shared__ FlatLds fl;
          if (j != staged) {  // uniform: the work item's job
            __syncthreads();
            stage_flat_layer(rc, fl, tid);
            __syncthreads();
            if (tid == 0) stage_flat_mins(rc, fl);
            __syncthreads();
            staged = j;
          }
          flp = &fl;
        }
        RefrFacet32 f32;
        {
          RefrFacet fr;
          load_refr_facet(k, fi, rc.flags, fr);
          to_f32(fr, f32);
        }
        if (tid == 0 && redo_base + gw >= rd.cap && rd.overflow) *rd.overflow = 1u;
        trace_loop<ACC, TT, true>(sh, k, w, live, f, rc.n_samples, split,
                                  [&](const TraceRec& t, PairOut& o, int) {
                                    const bool ok =
                                        (KIND == 2 && shallow)
                                            ? refracted_pair_mixed<(FS == FS_PLAIN ? kJointPlain : 4),
                                                                   RM, PAT>(f32, t, rc, o)
                                            : refracted_pair_mixed<NMAX, RM, PAT>(f32, t, rc, o,
                                                                                  flp);
                                    if ((rc.flags & SS_REDO_ALL) || !ok) {
                                      o.valid = false;
                                      o.pwr_invalid = 0.0;
                                      o.redo = true;
                                    }
                                  },
                                  rd, redo_base + gw);
      }
    }
    // Reconverge before the next work id: without this barrier the compiler
    // merges the ``tid == 0`` flag stores above with the ``tid == 0``
    // counter update at the loop top and lets the other lanes loop back to
    // the work-id read on their own, re-reading a stale id forever.
    __syncthreads();
  }
}

// Float64 pass over the refracted work items of one solver class (CODE)
// that a mixed batch_kernel launch flagged (the class's slice of the redo
// record), masked to the recorded lanes.
template <int TT, int CODE>
__global__ __launch_bounds__(kBlock) void batch_redo_kernel(BatchArgs ba, Redo rd) {
  constexpr int NMAX = class_nmax(CODE);
  __shared__ Shared<double, TT> sh;
  __shared__ long begins[kMaxJobs + 1];
  __shared__ unsigned long long sh_mask[kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const ss_job* __restrict__ jobs = ba.jobs;
  const int n_jobs = batch_n(ba);
  const long base = job_prefix<TT, CODE>(jobs, n_jobs, begins);
  long total = uniform_l(begins[n_jobs]);
  const long room = rd.cap - base;
  total = total < room ? total : (room > 0 ? room : 0);
  for_flagged_items(rd.flags + base, total, sh_mask, [&](long gw) {
    const int j = uniform_i(find_job(begins, n_jobs, gw));
    const ss_job& jb = jobs[j];
    if (uniform_i(job_code(jb)) != CODE) return;
    const long w = gw - begins[j];
    const KCommon k = job_common<TT>(jb);
    const long f = (w % k.n_ftiles) * kBlock + tid;
    const bool live = f < k.nf;
    const RefrConst rc = make_refr_const(jb.refracted, jb.iface_desc, jb.tables, jb.legs,
                                         jb.pattern_table, jb.constants);
    const bool split = (rc.flags & SS_SPLIT_SIDES) != 0;
    RefrFacet fr;
    load_refr_facet(k, live ? f : 0, rc.flags, fr);
    const unsigned long long* m = rd.masks + (base + gw) * (TT * 4);
    trace_loop<double, TT>(sh, k, w, live, f, rc.n_samples, split,
                           [&](const TraceRec& t, PairOut& o, int it) {
                             if (((m[it * 4 + wave] >> lane) & 1ULL) == 0ULL) {
                               o.redo = true;
                               return;
                             }
                             if (CODE == 2 && rc_shallow(jb.refracted))
                               refracted_pair_fp64<4>(fr, t, rc, o);
                             else
                               refracted_pair_fp64<NMAX>(fr, t, rc, o);
                           });
  });
}

// ------------------------------------------------------- range compression
constexpr int kTapChunk = 512;

__global__ __launch_bounds__(kBlock) void compress_kernel(const double2* __restrict__ in,
                                                          long n, const double2* __restrict__ pulse,
                                                          long m, double2* __restrict__ out) {
  __shared__ double2 sx[kBlock + kTapChunk];
  __shared__ double2 sp[kTapChunk];
  const long row = blockIdx.y;
  const long i0 = (long)blockIdx.x * kBlock;
  const int tid = threadIdx.x;
  const long ntaps = 2 * m + 1;
  double ar = 0.0, ai = 0.0;
  const double2* x = in + row * n;
  for (long j0 = 0; j0 < ntaps; j0 += kTapChunk) {
    int cnt = (int)((ntaps - j0) < kTapChunk ? (ntaps - j0) : kTapChunk);
    long lo = i0 + m - j0 - cnt + 1;
    for (int s = tid; s < kBlock + cnt - 1; s += kBlock) {
      long xi = lo + s;
      sx[s] = (xi >= 0 && xi < n) ? x[xi] : make_double2(0.0, 0.0);
    }
    for (int s = tid; s < cnt; s += kBlock) sp[s] = pulse[j0 + s];
    __syncthreads();
    const int base = tid + cnt - 1;
    for (int jj = 0; jj < cnt; ++jj) {
      double2 pj = sp[jj];
      double2 xv = sx[base - jj];
      ar = fma(pj.x, xv.x, ar);
      ar = fma(-pj.y, xv.y, ar);
      ai = fma(pj.x, xv.y, ai);
      ai = fma(pj.y, xv.x, ai);
    }
    __syncthreads();
  }
  if (i0 + tid < n) out[row * n + i0 + tid] = make_double2(ar, ai);
}

// FFT range compression (overlap-save).  The 'same' convolution
// out[i] = sum_j p[j] x[i + m - j] (2m+1 taps, zero lag at tap m) is the
// linear convolution y = p * x read from y[m].  One workgroup computes L =
// N - K + 1 consecutive outputs of one row (K = 2m+1): it loads the N input
// samples x[i0 + m - (K-1) ...] (zeros outside the row) into LDS, transforms
// them with a decimation-in-frequency FFT (natural order in, bit-reversed
// out), multiplies by the pulse spectrum (same bit-reversed order, scaled by
// 1/N), transforms back with a decimation-in-time inverse (bit-reversed in,
// natural out) and keeps the circular-convolution samples K-1 .. N-1, the
// ones free of wrap-around.  The whole N-point float64 transform lives in
// LDS (N <= 8192: 128 KiB); twiddles come from sincospi (one per thread
// per stage where the span allows, below), so no table is needed.  The
// kernels are launched with kFftThreads threads (the twiddle sharing
// relies on it).  This is the reference's own method
// (waveform.py:126-138 multiplies FFTs), blocked so any row length fits.
constexpr int kFftMax = 8192;
constexpr int kFftThreads = 512;

struct cd {
  double x, y;
};
__device__ inline cd cmul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// Butterfly twiddles: the butterflies a thread takes in a stage are b =
// tid + k kFftThreads, so for spans h <= kFftThreads (both powers of two)
// they share j = b & (h - 1) = tid & (h - 1) and one sincospi serves the
// whole stage; wider spans evaluate it per butterfly.  Same values either
// way.

// Forward DIF over a[0..n), n = 2^logn: natural order in, bit-reversed out.
__device__ inline void fft_dif(cd* a, int n, int logn) {
  for (int s = 0; s < logn; ++s) {
    const int lh = logn - 1 - s, h = 1 << lh;  // butterfly span
    const bool shared_j = h <= kFftThreads;
    double sn0 = 0.0, cs0 = 1.0;
    if (shared_j) sincospi(-(double)(threadIdx.x & (h - 1)) / (double)h, &sn0, &cs0);
    for (int b = threadIdx.x; b < (n >> 1); b += kFftThreads) {
      const int j = b & (h - 1);
      const int i0 = ((b >> lh) << (lh + 1)) + j, i1 = i0 + h;
      const cd u = a[i0], v = a[i1];
      double sn = sn0, cs = cs0;
      if (!shared_j) sincospi(-(double)j / (double)h, &sn, &cs);
      a[i0] = cd{u.x + v.x, u.y + v.y};
      a[i1] = cmul(cd{u.x - v.x, u.y - v.y}, cd{cs, sn});
    }
    __syncthreads();
  }
}

// Inverse DIT over a[0..n): bit-reversed in, natural order out (unscaled).
__device__ inline void ifft_dit(cd* a, int n, int logn) {
  for (int lh = 0; lh < logn; ++lh) {
    const int h = 1 << lh;
    const bool shared_j = h <= kFftThreads;
    double sn0 = 0.0, cs0 = 1.0;
    if (shared_j) sincospi((double)(threadIdx.x & (h - 1)) / (double)h, &sn0, &cs0);
    for (int b = threadIdx.x; b < (n >> 1); b += kFftThreads) {
      const int j = b & (h - 1);
      const int i0 = ((b >> lh) << (lh + 1)) + j, i1 = i0 + h;
      double sn = sn0, cs = cs0;
      if (!shared_j) sincospi((double)j / (double)h, &sn, &cs);
      const cd u = a[i0], t = cmul(a[i1], cd{cs, sn});
      a[i0] = cd{u.x + t.x, u.y + t.y};
      a[i1] = cd{u.x - t.x, u.y - t.y};
    }
    __syncthreads();
  }
}

// Pulse spectrum for compress_fft_kernel: DIF of the K taps zero-padded to
// n, scaled by 1/n (the inverse's normalisation), bit-reversed order.
__global__ __launch_bounds__(kFftThreads) void pulse_spectrum_kernel(const cd* __restrict__ pulse,
                                                                    long K, int n, int logn,
                                                                    cd* __restrict__ spec) {
  __shared__ cd a[kFftMax];
  for (int t = threadIdx.x; t < n; t += blockDim.x) a[t] = t < K ? pulse[t] : cd{0.0, 0.0};
  __syncthreads();
  fft_dif(a, n, logn);
  const double sc = 1.0 / (double)n;
  for (int t = threadIdx.x; t < n; t += blockDim.x) spec[t] = cd{a[t].x * sc, a[t].y * sc};
}

// grid (blocks per row, rows): block bx of row r produces out[i0, i0 + L).
// NCAP: LDS capacity in points (nfft <= NCAP); smaller capacities let more
// workgroups share a CU's 160 KiB of LDS.
template <int NCAP>
__global__ __launch_bounds__(kFftThreads) void compress_fft_kernel(const cd* __restrict__ in,
                                                                  long n, const cd* __restrict__ spec,
                                                                  long m, int nfft, int logn,
                                                                  cd* __restrict__ out) {
  __shared__ cd a[NCAP];
  const long K = 2 * m + 1, L = nfft - K + 1;
  const long row = blockIdx.y;
  const long i0 = (long)blockIdx.x * L;
  const long base = i0 + m - (K - 1);
  const cd* x = in + row * n;
  for (int t = threadIdx.x; t < nfft; t += blockDim.x) {
    const long xi = base + t;
    a[t] = (xi >= 0 && xi < n) ? x[xi] : cd{0.0, 0.0};
  }
  __syncthreads();
  fft_dif(a, nfft, logn);
  for (int t = threadIdx.x; t < nfft; t += blockDim.x) a[t] = cmul(a[t], spec[t]);
  __syncthreads();
  ifft_dit(a, nfft, logn);
  cd* o = out + row * n;
  for (long r = threadIdx.x; r < L; r += blockDim.x) {
    const long i = i0 + r;
    if (i < n) o[i] = a[K - 1 + r];
  }
}

// FFT size for rows of n samples and 2m+1 taps: the power of two in
// [2K, kFftMax] (at least 256) with the fewest butterflies per output
// sample, or 0 when the pulse is too long for an LDS-resident transform.
__host__ inline int compress_fft_size(long n, long m) {
  const long K = 2 * m + 1;
  int best = 0;
  double best_cost = 0.0;
  for (int N = 256; N <= kFftMax; N <<= 1) {
    const long L = N - K + 1;
    if (L < N / 4) continue;  // mostly wrap-around: too small
    const long blocks = (n + L - 1) / L;
    const double cost = (double)blocks * N * std::log2((double)N);
    if (best == 0 || cost < best_cost) {
      best = N;
      best_cost = cost;
    }
  }
  return best;
}

__global__ void fill_kernel(double* p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static int fill(double* p, long n, hipStream_t st) {
  if (n <= 0) return SS_OK;
  long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, 0.0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("fill: ") + hipGetErrorString(e));
  return SS_OK;
}

static int rough_mode(int flags) {
  if (!(flags & SS_ROUGH)) return 0;
  return (flags & SS_D_PHI_AREA) ? 1 : 2;
}

static int pattern_mode(int kind) {
  if (kind == SS_PAT_ISOTROPIC) return 0;
  if (kind == SS_PAT_TABULATED) return 5;
  return -1;
}

// Single launches: their tabulated instantiations stage the table in LDS
// (at most kPatLds samples); larger tables take the generic one.
static int launch_pattern_mode(int kind, int n) {
  if (kind == SS_PAT_TABULATED && n > kPatLds) return -1;
  return pattern_mode(kind);
}

static int check_pattern(int kind, int n, const double* table) {
  if (kind < SS_PAT_ISOTROPIC || kind > SS_PAT_TABULATED)
    return fail(SS_EINVAL, "unknown antenna pattern kind");
  if (kind == SS_PAT_TABULATED || kind == SS_PAT_ARRAY_TAPERED) {
    if (!table) return fail(SS_EINVAL, "pattern table required for this pattern kind");
    if (n < 2 || n > SS_MAX_PATTERN_TABLE)
      return fail(SS_ERANGE, "pattern table length out of range [2, 4096]");
  }
  return SS_OK;
}

static KCommon make_common(const double* facets, int64_t fs, int64_t nf,
                           const double* traces, int64_t ts, int64_t nt, double* field,
                           double* dropped, int32_t* pair_bins, int tt, int n_seg) {
  KCommon k;
  k.facets = facets;
  k.fs = fs;
  k.nf = nf;
  k.traces = traces;
  k.ts = ts;
  k.nt = nt;
  k.field = field;
  k.dropped = dropped;
  k.pair_bins = pair_bins;
  k.n_ftiles = (int)((nf + kBlock - 1) / kBlock);
  k.n_ttiles = (int)((nt + tt - 1) / tt);
  long total = (long)k.n_ftiles * k.n_ttiles;
  k.per_xcd = (total + 7) / 8;
  k.n_seg = n_seg;
  return k;
}

// ----------------------------------------------- along-track processing
// Boxcar along slow time (processing.py:143-238): out[k, c] = scale * sum_j
// in[k * stride + j, c], j < window; a complex (n, 2) interleaved array is
// 2 n real columns.  One thread per output column, rows by grid-stride.
__global__ __launch_bounds__(kBlock) void boxcar_kernel(const double* __restrict__ in, long cols,
                                                        long window, long stride, long n_out,
                                                        double scale, double* __restrict__ out) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= cols) return;
  for (long k = blockIdx.y; k < n_out; k += gridDim.y) {
    const double* p = in + k * stride * cols + c;
    double acc = 0.0;
    for (long j = 0; j < window; ++j) acc += p[j * cols];
    out[k * cols + c] = acc * scale;
  }
}

// Straight-track time-domain backprojection through air (processing.py:
// 270-348).  Workgroup = one output trace i x kBlock bins.  The aperture
// traces j with |s_j - s_i| <= half form one index range of the
// non-decreasing arc length s (found by bisection, tested exactly per
// trace); their offsets and window weights are staged in LDS a chunk at a
// time.  Per (j, bin): dr = sqrt(r0^2 + d^2) - r0, the field read at the
// migrated delay by linear interpolation (zero unless 0 <= i0 < nb - 1),
// times w exp(+2 j k dr).  Float64 tier: the reference's expressions and
// order, contraction off.  Mixed tier: float64 delay and phase cycles
// (2 k dr / 2 pi), float32 sin/cos, interpolation and sum.
constexpr int kFocusChunk = kBlock;

__device__ inline long lower_index(const double* s, long n, double v) {
  long lo = 0, hi = n;  // first index with s >= v
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (s[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void focus_kernel(const double* __restrict__ field, long T,
                                                       long nb, const double* __restrict__ twtt,
                                                       const double* __restrict__ s, double c,
                                                       double k, double half, int hann,
                                                       double* __restrict__ out) {
  __shared__ double sd[kFocusChunk];
  __shared__ double sw[kFocusChunk];
  __shared__ long range[2];
  const int tid = threadIdx.x;
  const long b = (long)blockIdx.x * kBlock + tid;
  const double t0 = twtt[0];
  const double dt = twtt[1] - twtt[0];
  const double cyc_per_m = 2.0 * k / kTwoPi;
  for (long i = blockIdx.y; i < T; i += gridDim.y) {
    const double si = s[i];
    if (tid == 0) {
      // one index of margin on each side: the exact test below decides
      const long lo = lower_index(s, T, si - half) - 1;
      const long hi = lower_index(s, T, si + half) + 2;  // past the last s <= si + half
      range[0] = lo < 0 ? 0 : lo;
      range[1] = hi > T ? T : hi;
    }
    __syncthreads();
    const long j0 = range[0], j1 = range[1];
    const double tw = b < nb ? twtt[b] : 0.0;
    double r0;
    {
#pragma clang fp contract(off)
      r0 = c * tw / 2.0;
    }
    const double r02 = r0 * r0;
    double are = 0.0, aim = 0.0;
    float fre = 0.0f, fim = 0.0f;
    for (long jc = j0; jc < j1; jc += kFocusChunk) {
      const int n = (int)((j1 - jc) < kFocusChunk ? (j1 - jc) : kFocusChunk);
      __syncthreads();  // the previous chunk is consumed
      if (tid < n) {
        const double d = s[jc + tid] - si;
        double w = -1.0;  // outside the aperture
        if (fabs(d) <= half) {
#pragma clang fp contract(off)
          w = hann ? 0.5 * (1.0 + cos(kPi * d / half)) : 1.0;
        }
        sd[tid] = d;
        sw[tid] = w;
      }
      __syncthreads();
      if (b < nb) {
        for (int q = 0; q < n; ++q) {
          const double w = sw[q];
          if (w < 0.0) continue;
          const double d = sd[q];
          const double* F = field + (jc + q) * nb * 2;
          if constexpr (PREC == SS_PREC_FP64) {
#pragma clang fp contract(off)
            const double dr = sqrt(r02 + d * d) - r0;
            const double tau = tw + 2.0 * dr / c;
            const double x = (tau - t0) / dt;
            const double fl = floor(x);
            if (!(fl >= 0.0 && fl < (double)(nb - 1))) continue;
            const long i0 = (long)fl;
            const double fr = x - fl;
            const double sre = F[2 * i0] * (1.0 - fr) + F[2 * i0 + 2] * fr;
            const double sim = F[2 * i0 + 1] * (1.0 - fr) + F[2 * i0 + 3] * fr;
            const double ph = (2.0 * k) * dr;
            const double cp = cos(ph), sp = sin(ph);
            const double wre = w * sre, wim = w * sim;
            are = are + (wre * cp - wim * sp);
            aim = aim + (wre * sp + wim * cp);
          } else {
            const double dr = sqrt(r02 + d * d) - r0;
            const double x = (tw + 2.0 * dr / c - t0) / dt;
            const double fl = floor(x);
            if (!(fl >= 0.0 && fl < (double)(nb - 1))) continue;
            const long i0 = (long)fl;
            const float fr = (float)(x - fl);
            const float sre = (float)F[2 * i0] * (1.0f - fr) + (float)F[2 * i0 + 2] * fr;
            const float sim = (float)F[2 * i0 + 1] * (1.0f - fr) + (float)F[2 * i0 + 3] * fr;
            const double cyc = dr * cyc_per_m;
            const float ph = (float)(cyc - rint(cyc));
            const float cp = f_cos_rev(ph), sp = f_sin_rev(ph);
            const float wre = (float)w * sre, wim = (float)w * sim;
            fre += wre * cp - wim * sp;
            fim += wre * sp + wim * cp;
          }
        }
      }
    }
    if (b < nb) {
      if constexpr (PREC == SS_PREC_FP64) {
        out[(i * nb + b) * 2] = are;
        out[(i * nb + b) * 2 + 1] = aim;
      } else {
        out[(i * nb + b) * 2] = (double)fre;
        out[(i * nb + b) * 2 + 1] = (double)fim;
      }
    }
    __syncthreads();  // range[] is rewritten for the next trace
  }
}

}  // namespace ss

using namespace ss;

// ---------------------------------------------------------- scene builder
// Facets of a regular vertex lattice built in HBM (the corner math of
// scene.py:158-176 / compare/brute_force_layered.py:44-71, float64 in the
// host restatement's operation order): a (ny+1, nx+1) height lattice at
// `spacing` centred on (x0, y0) (x along columns, y along rows).  Facet
// order: 16x16-cell tiles, tiles row-major, cells row-major in a tile, so a
// 256-facet workgroup tile is one compact patch (edge tiles narrower).
constexpr int kTileSide = 16;
constexpr int kStatCount = 10;  // sum A, sum A c (3), sum A n (3), sum c (3)

__host__ __device__ inline long tile_position(long r, long c, long ny, long nx) {
  const long tr = r / kTileSide, tc = c / kTileSide;
  const long h = (ny - tr * kTileSide) < kTileSide ? (ny - tr * kTileSide) : kTileSide;
  const long w = (nx - tc * kTileSide) < kTileSide ? (nx - tc * kTileSide) : kTileSide;
  return tr * kTileSide * nx + tc * kTileSide * h + (r % kTileSide) * w + (c % kTileSide);
}

__global__ __launch_bounds__(kBlock) void facets_from_heights_kernel(
    const double* __restrict__ heights, long ny, long nx, double spacing, double x0, double y0,
    double* __restrict__ facets, long fs, double* __restrict__ table,
    double* __restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double red[kStatCount][kBlock];
  const long n = ny * nx;
  const long cell = (long)blockIdx.x * kBlock + threadIdx.x;
  double st[kStatCount];
  for (int i = 0; i < kStatCount; ++i) st[i] = 0.0;
  if (cell < n) {
    const long r = cell / nx, c = cell - r * nx;
    const long nx1 = nx + 1, ny1 = ny + 1;
    const double hx = (double)(nx1 - 1) / 2.0, hy = (double)(ny1 - 1) / 2.0;
    const double xa = ((double)c - hx) * spacing + x0, xb = ((double)(c + 1) - hx) * spacing + x0;
    const double ya = ((double)r - hy) * spacing + y0, yb = ((double)(r + 1) - hy) * spacing + y0;
    const double v00[3] = {xa, ya, heights[r * nx1 + c]};
    const double v01[3] = {xb, ya, heights[r * nx1 + c + 1]};
    const double v10[3] = {xa, yb, heights[(r + 1) * nx1 + c]};
    const double v11[3] = {xb, yb, heights[(r + 1) * nx1 + c + 1]};
    double e1[3], e2[3], ce[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = ((v01[k] + v11[k]) - (v00[k] + v10[k])) / 2.0;
      e2[k] = ((v10[k] + v11[k]) - (v00[k] + v01[k])) / 2.0;
      ce[k] = (((v00[k] + v01[k]) + v10[k]) + v11[k]) / 4.0;
    }
    const double raw[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
    const double mag = sqrt((raw[0] * raw[0] + raw[1] * raw[1]) + raw[2] * raw[2]);
    double nr[3] = {raw[0] / mag, raw[1] / mag, raw[2] / mag};
    const double sg = nr[2] < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < 3; ++k) nr[k] = nr[k] * sg;
    const long pos = tile_position(r, c, ny, nx);
    for (int k = 0; k < 3; ++k) {
      facets[(SS_FROW_CX + k) * fs + pos] = ce[k];
      facets[(SS_FROW_NX + k) * fs + pos] = nr[k];
      facets[(SS_FROW_E1X + k) * fs + pos] = e1[k];
      facets[(SS_FROW_E2X + k) * fs + pos] = e2[k];
    }
    facets[SS_FROW_AREA * fs + pos] = mag;
    if (table) {
      for (int k = 0; k < 3; ++k) {
        table[cell * 6 + k] = ce[k];
        table[cell * 6 + 3 + k] = nr[k];
      }
    }
    st[0] = mag;
    for (int k = 0; k < 3; ++k) {
      st[1 + k] = mag * ce[k];
      st[4 + k] = mag * nr[k];
      st[7 + k] = ce[k];
    }
  }
  // block sums in a fixed tree order (deterministic)
  for (int i = 0; i < kStatCount; ++i) red[i][threadIdx.x] = st[i];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int i = 0; i < kStatCount; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + h];
    __syncthreads();
  }
  if ((int)threadIdx.x < kStatCount)
    partials[(long)blockIdx.x * kStatCount + threadIdx.x] = red[threadIdx.x][0];
}

// Sums the per-block partials in a fixed order: thread i strides the
// blocks, then a tree over the threads.
__global__ __launch_bounds__(kBlock) void facet_stats_kernel(const double* __restrict__ partials,
                                                             long n_blocks,
                                                             double* __restrict__ stats) {
  __shared__ double red[kStatCount][kBlock];
  double st[kStatCount];
  for (int i = 0; i < kStatCount; ++i) st[i] = 0.0;
  for (long b = threadIdx.x; b < n_blocks; b += kBlock)
    for (int i = 0; i < kStatCount; ++i) st[i] += partials[b * kStatCount + i];
  for (int i = 0; i < kStatCount; ++i) red[i][threadIdx.x] = st[i];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int i = 0; i < kStatCount; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + h];
    __syncthreads();
  }
  if ((int)threadIdx.x < kStatCount) stats[threadIdx.x] = red[threadIdx.x][0];
}

// ------------------------------------------------ seeded scene generators
// Keyed standard normals (roughness.py:207-233): lattice key (kr, kc) and a
// 64-bit stream base give keyed = base ^ kr ^ (kc * K), two splitmix64
// finaliser outputs (keyed, keyed ^ S2) -> u1, u2 in (0, 1) with 53 bits,
// then Box-Muller: radius sqrt(-2 log u1), angle 2 pi u2.  The speckle
// phasor is radius e^{i angle} / sqrt 2 rounded to complex64; the height
// noise of the Gaussian-ACF generator is its float64 real part x sqrt 2,
// radius cos(angle).  Integer steps are exact, so host and device agree to
// the ulps of log / cos / sin.
constexpr unsigned long long kKeyMul = 0xD2B74407B1CE6E93ull;
constexpr unsigned long long kStream2 = 0xCA5A826395121157ull;

__host__ __device__ inline unsigned long long splitmix_fin(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ inline void keyed_polar(unsigned long long base, long kr, long kc, double& rad,
                                   double& ang) {
#pragma clang fp contract(off)
  const unsigned long long keyed =
      base ^ (unsigned long long)kr ^ ((unsigned long long)kc * kKeyMul);
  const double u1 = ((double)(splitmix_fin(keyed) >> 11) + 0.5) * 0x1p-53;
  const double u2 = ((double)(splitmix_fin(keyed ^ kStream2) >> 11) + 0.5) * 0x1p-53;
  rad = sqrt(-2.0 * log(u1));
  ang = kTwoPi * u2;
}

// Row pass of the Gaussian-ACF height generator (compare/haynes.py:94-112):
// the circular convolution of keyed white noise with exp(-rho^2/a^2) is
// separable, exp(-(dy^2 + dx^2)/a^2) = gy(dy) gx(dx), so the reference's
// FFT product becomes a direct circular convolution along x (this kernel)
// then along y (gauss_cols_kernel).  One workgroup per lattice row: the
// row's noise and the taps gx[d], d = 0..nx1-1 (minimal-image distance),
// are staged in LDS; out[j] = sum_k w[k] gx[(j - k) mod nx1], split at k = j
// so no index wraps.
__global__ __launch_bounds__(kBlock) void gauss_rows_kernel(unsigned long long base, long nx1,
                                                            double spacing, double a2,
                                                            double* __restrict__ tmp) {
  HIP_DYNAMIC_SHARED(double, lds_g)
  double* w = lds_g;
  double* g = lds_g + nx1;
  const long row = blockIdx.x;
  for (long j = threadIdx.x; j < nx1; j += kBlock) {
    double rad, ang;
    keyed_polar(base, row, j, rad, ang);
    w[j] = rad * cos(ang);
    const long d = j < nx1 - j ? j : nx1 - j;
    const double c = (double)d * spacing;
    g[j] = exp(-(c * c) / a2);
  }
  __syncthreads();
  for (long j = threadIdx.x; j < nx1; j += kBlock) {
    double acc = 0.0;
    for (long k = 0; k <= j; ++k) acc = fma(w[k], g[j - k], acc);
    for (long k = j + 1; k < nx1; ++k) acc = fma(w[k], g[j - k + nx1], acc);
    tmp[row * nx1 + j] = acc;
  }
}

// Column pass: lanes are consecutive columns (coalesced loads), each lane
// keeps kGaussRows consecutive outputs in registers; the y taps are staged
// in LDS extended periodically to 2 ny1 + kGaussRows entries so that
// i0 + r - k + ny1 never wraps.  heights = mean + scale * sum.
constexpr int kGaussRows = 32;

__global__ __launch_bounds__(kBlock) void gauss_cols_kernel(const double* __restrict__ tmp,
                                                            long ny1, long nx1, double spacing,
                                                            double a2, double scale, double mean,
                                                            double* __restrict__ heights) {
  HIP_DYNAMIC_SHARED(double, lds_g)
  const long n_tap = 2 * ny1 + kGaussRows;
  for (long t = threadIdx.x; t < n_tap; t += kBlock) {
    const long m = t % ny1;
    const long d = m < ny1 - m ? m : ny1 - m;
    const double c = (double)d * spacing;
    lds_g[t] = exp(-(c * c) / a2);
  }
  __syncthreads();
  const long j = (long)blockIdx.x * kBlock + threadIdx.x;
  const long i0 = (long)blockIdx.y * kGaussRows;
  if (j >= nx1) return;  // no barrier below
  double acc[kGaussRows];
#pragma unroll
  for (int r = 0; r < kGaussRows; ++r) acc[r] = 0.0;
  for (long k = 0; k < ny1; ++k) {
    const double v = tmp[k * nx1 + j];
    const double* gk = lds_g + (i0 - k + ny1);  // gk[r] = gy[(i0 + r - k) mod ny1]
#pragma unroll
    for (int r = 0; r < kGaussRows; ++r) acc[r] = fma(v, gk[r], acc[r]);
  }
#pragma unroll
  for (int r = 0; r < kGaussRows; ++r)
    if (i0 + r < ny1) heights[(i0 + r) * nx1 + j] = mean + scale * acc[r];
}

// Keyed speckle phasors of a lattice's cells (roughness.py:207-233 with
// keys (kr0 + r, kc0 + c)), written to facet rows SS_FROW_PH_RE/IM at the
// cells' tile-order positions as float64 copies of the complex64 values.
__global__ __launch_bounds__(kBlock) void keyed_phasor_kernel(unsigned long long base, long ny,
                                                              long nx, long kr0, long kc0,
                                                              double* __restrict__ facets,
                                                              long fs) {
#pragma clang fp contract(off)
  const long cell = (long)blockIdx.x * kBlock + threadIdx.x;
  if (cell >= ny * nx) return;
  const long r = cell / nx, c = cell - r * nx;
  double rad, ang;
  keyed_polar(base, kr0 + r, kc0 + c, rad, ang);
  const double s2 = sqrt(2.0);
  const float re = (float)((rad * cos(ang)) / s2);
  const float im = (float)((rad * sin(ang)) / s2);
  const long pos = tile_position(r, c, ny, nx);
  facets[SS_FROW_PH_RE * fs + pos] = (double)re;
  facets[SS_FROW_PH_IM * fs + pos] = (double)im;
}

// ------------------------------------------------------ trace diagnostics
// Per trace, over one interface's facets (output.py:69-82,
// simulate.py:398-412): the horizontally nearest facet centre (nadir) and
// the earliest in-window exact arrival 2|p - c|/c (first return), ties to
// the lowest caller index as numpy's argmin.  kDiagTT traces per
// workgroup, each lane a strided subset of the facets, then a fixed-order
// LDS reduction per trace.
constexpr int kDiagTT = 8;

__device__ inline bool diag_better(double k, long i, double bk, long bi) {
  return k < bk || (k == bk && i < bi);
}

__global__ __launch_bounds__(kBlock) void trace_diag_kernel(
    const double* __restrict__ facets, long fs, long nf, const long* __restrict__ index,
    const double* __restrict__ traces, long ts, long nt, double t0, double t_end, double c,
    long* __restrict__ nadir_index, double* __restrict__ nadir_z,
    long* __restrict__ first_index, double* __restrict__ first_twtt,
    double* __restrict__ first_xyz) {
#pragma clang fp contract(off)
  __shared__ double sk[kBlock];
  __shared__ long si[kBlock], sp[kBlock];
  const int tid = threadIdx.x;
  const long t_base = (long)blockIdx.x * kDiagTT;
  const int nT = (int)((nt - t_base) < kDiagTT ? (nt - t_base) : kDiagTT);
  double px[kDiagTT], py[kDiagTT], pz[kDiagTT];
  double bn[kDiagTT], bf[kDiagTT];
  long bni[kDiagTT], bnp[kDiagTT], bfi[kDiagTT], bfp[kDiagTT];
  const double inf = __builtin_inf();
  for (int t = 0; t < kDiagTT; ++t) {
    const long tt = t_base + (t < nT ? t : 0);
    px[t] = traces[(SS_TROW_PX + 0) * ts + tt];
    py[t] = traces[(SS_TROW_PX + 1) * ts + tt];
    pz[t] = traces[(SS_TROW_PX + 2) * ts + tt];
    bn[t] = bf[t] = inf;
    bni[t] = bfi[t] = LONG_MAX;
    bnp[t] = bfp[t] = -1;
  }
  for (long f = tid; f < nf; f += kBlock) {
    const double cx = facets[(SS_FROW_CX + 0) * fs + f];
    const double cy = facets[(SS_FROW_CX + 1) * fs + f];
    const double cz = facets[(SS_FROW_CX + 2) * fs + f];
    const long oi = index ? index[f] : f;
    for (int t = 0; t < kDiagTT; ++t) {
      const double dx = cx - px[t], dy = cy - py[t], dz = cz - pz[t];
      const double h2 = dx * dx + dy * dy;
      if (diag_better(h2, oi, bn[t], bni[t])) {
        bn[t] = h2;
        bni[t] = oi;
        bnp[t] = f;
      }
      const double ft = 2.0 * sqrt(h2 + dz * dz) / c;
      if (ft >= t0 && ft < t_end && diag_better(ft, oi, bf[t], bfi[t])) {
        bf[t] = ft;
        bfi[t] = oi;
        bfp[t] = f;
      }
    }
  }
  for (int t = 0; t < nT; ++t) {
    for (int q = 0; q < 2; ++q) {
      sk[tid] = q == 0 ? bn[t] : bf[t];
      si[tid] = q == 0 ? bni[t] : bfi[t];
      sp[tid] = q == 0 ? bnp[t] : bfp[t];
      __syncthreads();
      for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (tid < h && diag_better(sk[tid + h], si[tid + h], sk[tid], si[tid])) {
          sk[tid] = sk[tid + h];
          si[tid] = si[tid + h];
          sp[tid] = sp[tid + h];
        }
        __syncthreads();
      }
      if (tid == 0) {
        const long tt = t_base + t;
        const long pos = sp[0];
        if (q == 0) {
          nadir_index[tt] = pos >= 0 ? si[0] : -1;
          nadir_z[tt] = pos >= 0 ? facets[(SS_FROW_CX + 2) * fs + pos] : __builtin_nan("");
        } else {
          first_index[tt] = pos >= 0 ? si[0] : -1;
          first_twtt[tt] = pos >= 0 ? sk[0] : inf;
          for (int k = 0; k < 3; ++k)
            first_xyz[tt * 3 + k] = pos >= 0 ? facets[(SS_FROW_CX + k) * fs + pos]
                                             : __builtin_nan("");
        }
      }
      __syncthreads();  // the LDS arrays are rewritten for the next reduction
    }
  }
}

extern "C" {

const char* ss_version(void) { return "soundersim-hip 0.1 (gfx950)"; }

const char* ss_last_error(void) { return g_last_error.c_str(); }

int ss_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int ss_fill_constants(double* out) {
  if (!out) return fail(SS_EINVAL, "null output");
  // Weideman (1994): L = sqrt(N/sqrt(2)), samples of exp(-t^2)(L^2 + t^2) at
  // t = L tan(k pi / 2M), M = 2N, FFT -> polynomial coefficients.
  const int N = 32, M = 2 * N;
  const double L = sqrt(N / sqrt(2.0));
  double f[2 * M];
  f[0] = 0.0;
  for (int i = 1; i < 2 * M; ++i) {
    int kk = -M + i;  // k = -M+1 .. M-1
    double t = L * tan(kPi * kk / (2.0 * M));
    f[i] = exp(-t * t) * (L * L + t * t);
  }
  // fftshift then real part of DFT / (2M)
  double g[2 * M];
  for (int i = 0; i < 2 * M; ++i) g[i] = f[(i + M) % (2 * M)];
  double a[2 * M];
  for (int j = 0; j < 2 * M; ++j) {
    double acc = 0.0;
    for (int i = 0; i < 2 * M; ++i) acc += g[i] * cos(-2.0 * kPi * (double)i * j / (2.0 * M));
    a[j] = acc / (2.0 * M);
  }
  out[SS_CONST_WEIDEMAN_L] = L;
  for (int i = 0; i < N; ++i) out[SS_CONST_WEIDEMAN_A + i] = a[N - i];  // a[1..N] reversed
  for (int m = 0; m <= SS_MAX_TERMS; ++m) out[SS_CONST_LGAMMA + m] = lgamma((double)m + 1.0);
  g_last_error.clear();
  return SS_OK;
}

int ss_surface_field(int precision, const ss_surface_params* prm, const double* facets,
                     int64_t facet_stride, int64_t n_facets, const double* traces,
                     int64_t trace_stride, int64_t n_traces, const double* pattern_table,
                     const double* constants, double* field, double* dropped,
                     int32_t* pair_bins, void* stream) {
  if (!prm) return fail(SS_EINVAL, "null params");
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (n_facets < 0 || n_traces < 0) return fail(SS_EINVAL, "negative sizes");
  if (prm->n_samples <= 0) return fail(SS_EINVAL, "n_samples must be > 0");
  if (!(prm->dt > 0.0) || !(prm->c > 0.0)) return fail(SS_EINVAL, "dt and c must be > 0");
  if (n_traces > 0 && (!field || !dropped || !traces))
    return fail(SS_EINVAL, "null trace/output pointer");
  if (n_facets > 0 && !facets) return fail(SS_EINVAL, "null facets");
  if (facet_stride < n_facets || trace_stride < n_traces)
    return fail(SS_EINVAL, "stride smaller than count");
  if ((prm->flags & SS_ROUGH)) {
    if (!constants) return fail(SS_EINVAL, "roughness needs the constants table");
    if (prm->n_terms < 0 || prm->n_terms > SS_MAX_TERMS)
      return fail(SS_ERANGE, "n_terms out of range [0, 300]");
    if (prm->acf == SS_ACF_EXPONENTIAL && !(prm->flags & SS_D_PHI_AREA))
      return fail(SS_EINVAL, "roughness acf='exponential' requires the area-only D_Phi");
  }
  int rcp = check_pattern(prm->pattern, prm->pat_n, pattern_table);
  if (rcp) return rcp;
  const bool split = (prm->flags & SS_SPLIT_SIDES) != 0;
  const int n_seg = (split ? 2 : 1) * prm->n_samples;
  if ((long)n_seg * 2 > (long)INT_MAX) return fail(SS_ERANGE, "n_samples too large");
  hipStream_t st = (hipStream_t)stream;
  int r = fill(field, (long)n_traces * n_seg * 2, st);
  if (r) return r;
  r = fill(dropped, n_traces, st);
  if (r) return r;
  if (n_traces == 0 || n_facets == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  KCommon k = make_common(facets, facet_stride, n_facets, traces, trace_stride, n_traces,
                          field, dropped, pair_bins, kTTSurf, n_seg);
  const SurfaceConst sc = make_surface_const(*prm, pattern_table, constants);
  dim3 grid((unsigned)(8 * k.per_xcd)), block(kBlock);
  const int rm = rough_mode(prm->flags), pm = launch_pattern_mode(prm->pattern, prm->pat_n);
#define SS_SURF(P, R, A)                                                                      \
  if (P == SS_PREC_MIXED && !(prm->flags & ~kSurfPlainRuntimeFlags))                          \
    hipLaunchKernelGGL((surface_kernel<P, R, A, FS_SURF_PLAIN>), grid, block, 0, st, k, sc);  \
  else                                                                                        \
    hipLaunchKernelGGL((surface_kernel<P, R, A>), grid, block, 0, st, k, sc)
#define SS_SURF_PA(P, R) \
  if (pm == 0) SS_SURF(P, R, 0); else if (pm == 5) SS_SURF(P, R, 5); else SS_SURF(P, R, -1)
#define SS_SURF_P(P) \
  if (rm == 0) { SS_SURF_PA(P, 0); } else if (rm == 1) { SS_SURF_PA(P, 1); } \
  else { SS_SURF_PA(P, 2); }
  if (precision == SS_PREC_FP64) {
    SS_SURF_P(SS_PREC_FP64);
  } else {
    SS_SURF_P(SS_PREC_MIXED);
  }
#undef SS_SURF_P
#undef SS_SURF_PA
#undef SS_SURF
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("surface launch: ") + hipGetErrorString(e));
  if (precision == SS_PREC_FP64) {
    long n = (long)n_traces * n_seg;
    hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, field,
                       traces, (long)trace_stride, (long)n_traces, n_seg, prm->k);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SS_EHIP, std::string("fold launch: ") + hipGetErrorString(e));
  }
  g_last_error.clear();
  return SS_OK;
}

int ss_incoherent_power(int precision, const ss_incoherent_params* prm, const double* facets,
                        int64_t facet_stride, int64_t n_facets, const double* traces,
                        int64_t trace_stride, int64_t n_traces, const double* pattern_table,
                        double* power, double* dropped, int32_t* pair_bins, void* stream) {
  if (!prm) return fail(SS_EINVAL, "null params");
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (n_facets < 0 || n_traces < 0) return fail(SS_EINVAL, "negative sizes");
  if (prm->n_samples <= 0) return fail(SS_EINVAL, "n_samples must be > 0");
  if (!(prm->dt > 0.0) || !(prm->c > 0.0)) return fail(SS_EINVAL, "dt and c must be > 0");
  if (prm->flags & ~SS_SPLIT_SIDES)
    return fail(SS_EINVAL, "the incoherent kernel takes no flag but SS_SPLIT_SIDES");
  if (n_traces > 0 && (!power || !dropped || !traces))
    return fail(SS_EINVAL, "null trace/output pointer");
  if (n_facets > 0 && !facets) return fail(SS_EINVAL, "null facets");
  if (facet_stride < n_facets || trace_stride < n_traces)
    return fail(SS_EINVAL, "stride smaller than count");
  int rcp = check_pattern(prm->pattern, prm->pat_n, pattern_table);
  if (rcp) return rcp;
  const bool split = (prm->flags & SS_SPLIT_SIDES) != 0;
  const int n_seg = (split ? 2 : 1) * prm->n_samples;
  if ((long)n_seg * 2 > (long)INT_MAX) return fail(SS_ERANGE, "n_samples too large");
  hipStream_t st = (hipStream_t)stream;
  int r = fill(power, (long)n_traces * n_seg, st);
  if (r) return r;
  r = fill(dropped, n_traces, st);
  if (r) return r;
  if (n_traces == 0 || n_facets == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  ss_surface_params sp = {};
  sp.t0 = prm->t0;
  sp.dt = prm->dt;
  sp.c = prm->c;
  sp.pat_a = prm->pat_a;
  for (int i = 0; i < 4; ++i) sp.pat_b[i] = prm->pat_b[i];
  sp.n_samples = prm->n_samples;
  sp.flags = prm->flags;
  sp.pattern = prm->pattern;
  sp.pat_n = prm->pat_n;
  KCommon k = make_common(facets, facet_stride, n_facets, traces, trace_stride, n_traces,
                          power, dropped, pair_bins, kTTSurf, n_seg);
  const int fpl = precision == SS_PREC_MIXED ? incoherent_fpl<SS_PREC_MIXED>()
                                             : incoherent_fpl<SS_PREC_FP64>();
  k.n_ftiles = (int)((n_facets + (long)kBlock * fpl - 1) / ((long)kBlock * fpl));
  k.per_xcd = ((long)k.n_ftiles * k.n_ttiles + 7) / 8;
  const SurfaceConst sc = make_surface_const(sp, pattern_table, nullptr);
  dim3 grid((unsigned)(8 * k.per_xcd)), block(kBlock);
  const int pm = launch_pattern_mode(prm->pattern, prm->pat_n);
#define SS_INC(P, A) \
  hipLaunchKernelGGL((incoherent_kernel<P, A>), grid, block, 0, st, k, sc)
#define SS_INC_P(P) \
  if (pm == 0) SS_INC(P, 0); else if (pm == 5) SS_INC(P, 5); else SS_INC(P, -1)
  if (precision == SS_PREC_FP64) {
    SS_INC_P(SS_PREC_FP64);
  } else {
    SS_INC_P(SS_PREC_MIXED);
  }
#undef SS_INC_P
#undef SS_INC
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SS_EHIP, std::string("incoherent launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_refracted_field(int precision, const ss_refracted_params* prm, const double* facets,
                       int64_t facet_stride, int64_t n_facets, const double* traces,
                       int64_t trace_stride, int64_t n_traces, const double* iface_desc,
                       const double* tables, const double* legs, const double* pattern_table,
                       const double* constants, double* out, double* dropped,
                       int32_t* pair_bins, void* workspace, int64_t workspace_bytes,
                       void* stream) {
  if (!prm) return fail(SS_EINVAL, "null params");
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (n_facets < 0 || n_traces < 0) return fail(SS_EINVAL, "negative sizes");
  if (prm->n_samples <= 0) return fail(SS_EINVAL, "n_samples must be > 0");
  if (!(prm->dt > 0.0) || !(prm->c > 0.0)) return fail(SS_EINVAL, "dt and c must be > 0");
  if (prm->n_crossed < 1 || prm->n_crossed > SS_MAX_CROSSED)
    return fail(SS_ERANGE, "n_crossed out of range [1, 64]");
  if (!iface_desc || !tables || !legs) return fail(SS_EINVAL, "null interface data");
  if (prm->refraction < 0 || prm->refraction > 2)
    return fail(SS_EINVAL, "refraction must be 0 (sequential), 1 (joint) or 2 (flat-stack joint)");
  if (prm->refraction != 0) {
    if (prm->refraction == 1 && prm->n_crossed > SS_MAX_JOINT)
      return fail(SS_ERANGE, "joint stack deeper than SS_MAX_JOINT (32)");
    if (prm->k_pad < 0 || prm->k_pad >= prm->n_crossed)
      return fail(SS_EINVAL, "k_pad must be in [0, n_crossed)");
    if (prm->joint_newton < 0 || prm->joint_backtrack < 1 || prm->joint_newton > 64 ||
        prm->joint_backtrack > 32)
      return fail(SS_ERANGE, "joint budgets out of range (newton 0..64, backtrack 1..32)");
  }
  if (n_traces > 0 && (!out || !dropped || !traces))
    return fail(SS_EINVAL, "null trace/output pointer");
  if (n_facets > 0 && !facets) return fail(SS_EINVAL, "null facets");
  if (facet_stride < n_facets || trace_stride < n_traces)
    return fail(SS_EINVAL, "stride smaller than count");
  const bool coherent = (prm->flags & SS_COHERENT) != 0;
  const int feat = SS_ROUGH | SS_CROSS_ROUGH | SS_TAPER | SS_D_PHI_AREA | SS_GAMMA_FACET | SS_DIFFUSE;
  if (!coherent && (prm->flags & feat))
    return fail(SS_EINVAL, "roughness, grazing fix, per-facet gamma and diffuse require coherent mode");
  if (prm->flags & SS_ROUGH) {
    if (!constants) return fail(SS_EINVAL, "roughness needs the constants table");
    if (prm->n_terms < 0 || prm->n_terms > SS_MAX_TERMS)
      return fail(SS_ERANGE, "n_terms out of range [0, 300]");
    if (prm->acf == SS_ACF_EXPONENTIAL && !(prm->flags & SS_D_PHI_AREA))
      return fail(SS_EINVAL, "roughness acf='exponential' requires the area-only D_Phi");
  }
  int rcp = check_pattern(prm->pattern, prm->pat_n, pattern_table);
  if (rcp) return rcp;
  const int64_t need = ss_refracted_workspace_bytes(precision, n_facets, n_traces);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(SS_EINVAL, "workspace smaller than ss_refracted_workspace_bytes()");
  if (need > 0 && (reinterpret_cast<uintptr_t>(workspace) & 7) != 0)
    return fail(SS_EINVAL, "workspace must be 8-byte aligned");
  const bool split = (prm->flags & SS_SPLIT_SIDES) != 0;
  const int n_seg = (split ? 2 : 1) * prm->n_samples;
  hipStream_t st = (hipStream_t)stream;
  const long out_len = (long)n_traces * n_seg * 2;
  int r = fill(out, out_len, st);
  if (r) return r;
  r = fill(dropped, n_traces, st);
  if (r) return r;
  if (n_traces == 0 || n_facets == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  KCommon k = make_common(facets, facet_stride, n_facets, traces, trace_stride, n_traces, out,
                          dropped, pair_bins, kTTRefr, n_seg);
  const RefrConst rc = make_refr_const(*prm, iface_desc, tables, legs, pattern_table,
                                       constants);
  dim3 grid((unsigned)(8 * k.per_xcd)), block(kBlock);
  const bool fp64 = precision == SS_PREC_FP64;
  const int rm = rough_mode(prm->flags), pm = launch_pattern_mode(prm->pattern, prm->pat_n);
  const bool plain = plain_profile(*prm);
  const long n_items = (long)k.n_ftiles * k.n_ttiles;
  const Redo rd = fp64 ? Redo{} : redo_view(workspace, n_items);
  // Redo pass: each block sweeps chunks of kBlock item flags.
  const long rchunks = (n_items + kBlock - 1) / kBlock;
  dim3 rgrid((unsigned)(rchunks < 1024 ? rchunks : 1024));
#define SS_REFR(P, N, R, A) \
  hipLaunchKernelGGL((refracted_kernel<P, N, R, A>), grid, block, 0, st, k, rc, rd)
#define SS_REDO(N, R, A) \
  hipLaunchKernelGGL((redo_kernel<N, R, A>), rgrid, block, 0, st, k, rc, rd)
#define SS_BOTH(N, R, A) \
  if (fp64) { SS_REFR(SS_PREC_FP64, N, R, A); } \
  else { SS_REFR(SS_PREC_MIXED, N, R, A); SS_REDO(N, R, A); }
#define SS_SEQ(R, A)                                                                          \
  if (!fp64 && plain) {                                                                       \
    hipLaunchKernelGGL((refracted_kernel<SS_PREC_MIXED, 0, R, A, FS_PLAIN>), grid, block, 0,  \
                       st, k, rc, rd);                                                        \
    SS_REDO(0, R, A);                                                                         \
  } else {                                                                                    \
    SS_BOTH(0, R, A);                                                                         \
  }
#define SS_REFR_PA(R) \
  if (pm == 0) { SS_SEQ(R, 0); } else if (pm == 5) { SS_SEQ(R, 5); } \
  else { SS_SEQ(R, -1); }
#define SS_FLAT(R, A)                                                                         \
  if (fp64) {                                                                                 \
    SS_REFR(SS_PREC_FP64, -1, R, A);                                                          \
  } else {                                                                                    \
    if (plain_flags(*prm))                                                                    \
      hipLaunchKernelGGL((flat_kernel<SS_PREC_MIXED, R, A, FS_PLAIN>), grid, block, 0, st, k,   \
                         rc, rd);                                                             \
    else                                                                                      \
      hipLaunchKernelGGL((flat_kernel<SS_PREC_MIXED, R, A>), grid, block, 0, st, k, rc, rd);  \
    SS_REDO(-1, R, A);                                                                        \
  }
  if (prm->refraction == 2) {
    // flat stacks: the multilayer scenes' smooth interfaces under a
    // tabulated (or no) pattern get their own instantiations
    if (rm == 0 && pm == 5) {
      SS_FLAT(0, 5);
    } else if (rm == 0 && pm == 0) {
      SS_FLAT(0, 0);
    } else {
      SS_FLAT(-1, -1);
    }
  } else if (!rc.joint) {
    if (rm == 0) {
      SS_REFR_PA(0);
    } else if (rm == 1) {
      SS_REFR_PA(1);
    } else {
      SS_REFR_PA(2);
    }
  } else if (prm->n_crossed <= 4) {
    // smooth interfaces under an isotropic or tabulated pattern, plain
    // flags: the joint solve built for the stack's own depth bound (two
    // and three deep at 2 waves per SIMD, as the batched classes)
    if (!fp64 && plain_flags(*prm) && rm == 0 && (pm == 0 || pm == 5)) {
#define SS_JPLAIN(N, A)                                                                      \
  hipLaunchKernelGGL((refracted_kernel<SS_PREC_MIXED, N, 0, A, FS_PLAIN>), grid, block, 0,   \
                     st, k, rc, rd)
#define SS_JPLAIN_A(N) \
  if (pm == 0) { SS_JPLAIN(N, 0); } else { SS_JPLAIN(N, 5); }
      if (prm->n_crossed <= 2) {
        SS_JPLAIN_A(2);
      } else if (prm->n_crossed <= 3) {
        SS_JPLAIN_A(3);
      } else {
        SS_JPLAIN_A(4);
      }
#undef SS_JPLAIN_A
#undef SS_JPLAIN
      SS_REDO(4, -1, -1);
    } else {
      SS_BOTH(4, -1, -1);
    }
  } else if (prm->n_crossed <= 8) {
    SS_BOTH(8, -1, -1);
  } else {
    SS_BOTH(SS_MAX_JOINT, -1, -1);
  }
#undef SS_FLAT
#undef SS_REFR_PA
#undef SS_SEQ
#undef SS_BOTH
#undef SS_REDO
#undef SS_REFR
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("refracted launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

}  // extern "C"

namespace ss {
// Workspace of ss_refracted_field_noalloc: one array per GPU, part of the
// code object, so it exists from library load on and is never allocated or
// freed.  In host code ``scratch`` names a host-side stand-in; the device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[SS_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

// Device address of the current device's scratch array.
static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SS_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(SS_ERANGE, "device index beyond 64");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return SS_OK;
}
}  // namespace ss

extern "C" {

int ss_refracted_field_noalloc(int precision, const ss_refracted_params* prm,
                               const double* facets, int64_t facet_stride, int64_t n_facets,
                               const double* traces, int64_t trace_stride, int64_t n_traces,
                               const double* iface_desc, const double* tables,
                               const double* legs, const double* pattern_table,
                               const double* constants, double* out, double* dropped,
                               int32_t* pair_bins, void* stream) {
  const int64_t need = ss_refracted_workspace_bytes(precision, n_facets, n_traces);
  if (need > SS_NOALLOC_SCRATCH_BYTES)
    return fail(SS_ERANGE, "ss_refracted_field_noalloc: the call needs " + std::to_string(need) +
                               " workspace bytes, more than SS_NOALLOC_SCRATCH_BYTES (" +
                               std::to_string(SS_NOALLOC_SCRATCH_BYTES) +
                               "); use ss_refracted_field with a workspace");
  void* ws = nullptr;
  if (need > 0) {
    const int r = scratch_address(&ws);
    if (r) return r;
  }
  return ss_refracted_field(precision, prm, facets, facet_stride, n_facets, traces,
                            trace_stride, n_traces, iface_desc, tables, legs, pattern_table,
                            constants, out, dropped, pair_bins, ws,
                            need > 0 ? SS_NOALLOC_SCRATCH_BYTES : 0, stream);
}

int64_t ss_refracted_workspace_bytes(int precision, int64_t n_facets, int64_t n_traces) {
  if (precision != SS_PREC_MIXED || n_facets <= 0 || n_traces <= 0) return 0;
  const long n_items = ((n_facets + kBlock - 1) / kBlock) * ((n_traces + kTTRefr - 1) / kTTRefr);
  return redo_bytes(n_items, kTTRefr);
}

int64_t ss_sizeof_job(void) { return (int64_t)sizeof(ss_job); }

static int validate_job(const ss_job& j, int idx) {
  std::string pre = "job " + std::to_string(idx) + ": ";
  if (j.kind != 0 && j.kind != 1) return fail(SS_EINVAL, pre + "kind must be 0 or 1");
  if (j.n_facets < 0 || j.n_traces < 0) return fail(SS_EINVAL, pre + "negative sizes");
  if (j.facet_stride < j.n_facets || j.trace_stride < j.n_traces)
    return fail(SS_EINVAL, pre + "stride smaller than count");
  if (j.n_traces > 0 && (!j.traces || !j.field || !j.dropped))
    return fail(SS_EINVAL, pre + "null trace/output pointer");
  if (j.n_facets > 0 && !j.facets) return fail(SS_EINVAL, pre + "null facets");
  int n_samples, flags, pattern, pat_n, n_terms, acf;
  double dt, c;
  if (j.kind == 0) {
    n_samples = j.surface.n_samples;
    flags = j.surface.flags;
    pattern = j.surface.pattern;
    pat_n = j.surface.pat_n;
    n_terms = j.surface.n_terms;
    acf = j.surface.acf;
    dt = j.surface.dt;
    c = j.surface.c;
  } else {
    const ss_refracted_params& r = j.refracted;
    n_samples = r.n_samples;
    flags = r.flags;
    pattern = r.pattern;
    pat_n = r.pat_n;
    n_terms = r.n_terms;
    acf = r.acf;
    dt = r.dt;
    c = r.c;
    if (r.refraction < 0 || r.refraction > 2)
      return fail(SS_EINVAL, pre + "refraction must be 0 (sequential), 1 (joint) or 2 "
                                   "(flat-stack joint)");
    if (r.n_crossed < 1 || r.n_crossed > SS_MAX_CROSSED)
      return fail(SS_ERANGE, pre + "n_crossed out of range [1, 64]");
    if (r.refraction != 0) {
      if (r.refraction == 1 && r.n_crossed > SS_MAX_JOINT)
        return fail(SS_ERANGE, pre + "joint stack deeper than SS_MAX_JOINT (32)");
      if (r.k_pad < 0 || r.k_pad >= r.n_crossed)
        return fail(SS_EINVAL, pre + "k_pad must be in [0, n_crossed)");
      if (r.joint_newton < 0 || r.joint_backtrack < 1 || r.joint_newton > 64 ||
          r.joint_backtrack > 32)
        return fail(SS_ERANGE, pre + "joint budgets out of range (newton 0..64, backtrack 1..32)");
    }
    if (!j.iface_desc || !j.tables || !j.legs)
      return fail(SS_EINVAL, pre + "null interface data");
  }
  if (n_samples <= 0) return fail(SS_EINVAL, pre + "n_samples must be > 0");
  if (!(dt > 0.0) || !(c > 0.0)) return fail(SS_EINVAL, pre + "dt and c must be > 0");
  if (flags & SS_ROUGH) {
    if (!j.constants) return fail(SS_EINVAL, pre + "roughness needs the constants table");
    if (n_terms < 0 || n_terms > SS_MAX_TERMS) return fail(SS_ERANGE, pre + "n_terms range");
    if (acf == SS_ACF_EXPONENTIAL && !(flags & SS_D_PHI_AREA))
      return fail(SS_EINVAL, pre + "exponential ACF requires the area-only D_Phi");
  }
  return check_pattern(pattern, pat_n, j.pattern_table);
}

static int64_t batch_items(const ss_job* jobs_host, int n_jobs, int trace_tile) {
  int64_t total = 0;
  for (int i = 0; i < n_jobs; ++i) total += job_work(jobs_host[i], trace_tile);
  return total;
}

// A job the plain flag profile covers: a surface job without split sides,
// interpolated bins or per-facet gamma; a coherent refracted job without
// split sides, diffuse term, per-facet gamma or forced redo, over one
// crossing (sequential chain) or a stack of at most kJointPlain (joint
// solve).
static bool batch_plain_job(const ss_job& j) {
  if (j.kind == 0) return !(j.surface.flags & ~kSurfPlainRuntimeFlags);
  const int code = job_code(j);
  if (code == 1) return plain_profile(j.refracted);
  if (code == 2) return plain_flags(j.refracted) && j.refracted.n_crossed <= kJointPlain;
  return plain_flags(j.refracted);  // flat stack, two-deep joint solve
}

// Feature class of a job mix: 0 = no roughness, isotropic antenna and every
// job plain (batch_plain_job; the mixed launches of class 0 fold those
// flags at compile time and build the joint solve for stacks of at most
// kJointPlain interfaces only); 1 = at most area-only D_Phi and a tabulated pattern; 2 =
// anything.  Each class's kernels run the jobs of the classes below it.
static int batch_class(const ss_job* jobs_host, int n_jobs) {
  int cls = 0;
  for (int i = 0; i < n_jobs; ++i) {
    const ss_job& j = jobs_host[i];
    const int fl = j.kind == 0 ? j.surface.flags : j.refracted.flags;
    const int pk = j.kind == 0 ? j.surface.pattern : j.refracted.pattern;
    const int rm = rough_mode(fl), pm = pattern_mode(pk);
    int c = (rm == 0 && pm == 0) ? 0 : ((rm <= 1 && pm != -1) ? 1 : 2);
    if (c == 0 && !batch_plain_job(j)) c = 1;
    cls = c > cls ? c : cls;
  }
  return cls;
}

static int device_cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return cus > 0 ? cus : 1;
}

static int default_blocks(int n_blocks) { return n_blocks > 0 ? n_blocks : 4 * device_cus(); }

// The launches of one batched step on a job table already in device memory:
// zero the outputs, reset the per-class work counters and the overflow
// word, one persistent launch per job class (grid[code] workgroups; 0 =
// skip), the float64 redo sweep of each refracted class (mixed; rblocks[code]
// workgroups) and the fold-back of float64 surface jobs.  Nothing here
// depends on the table's contents on the host side when grids are fixed,
// so the sequence can be captured once and replayed with new tables.  The
// surface and sequential classes take the feature class's instantiation;
// the joint classes the generic one.
static void launch_batch(bool fp64, const BatchArgs& ba, void* workspace, const Redo& rd,
                         int tt, const unsigned grid[kJobCodes],
                         const unsigned rblocks[kJobCodes], int cls, bool fold,
                         hipStream_t st) {
  unsigned long long* hdr = static_cast<unsigned long long*>(workspace);
  const dim3 block(kBlock);
  const unsigned cus = (unsigned)device_cus();
  hipLaunchKernelGGL(batch_zero_kernel, dim3(cus, 8), dim3(256), 0, st, ba);
  hipLaunchKernelGGL(reset_workspace_kernel, dim3(1), dim3(64), 0, st, hdr);
  for (int code = 0; code < kJobCodes; ++code) {
    if (grid[code] == 0) continue;
    const dim3 g(grid[code]);
    const dim3 rg(rblocks[code]);
    unsigned long long* ctr = hdr + code;
#define SS_BATCH_K(P, T, R, A)                                                                \
  if (code == 0)                                                                              \
    hipLaunchKernelGGL((batch_kernel<P, T, R, A, 0>), g, block, 0, st, ba, ctr, rd);          \
  else                                                                                        \
    hipLaunchKernelGGL((batch_kernel<P, T, R, A, 1>), g, block, 0, st, ba, ctr, rd)
#define SS_BATCH_PLAIN(T)                                                                     \
  if (code == 0)                                                                              \
    hipLaunchKernelGGL((batch_kernel<SS_PREC_MIXED, T, 0, 0, 0, FS_PLAIN>), g, block, 0, st,  \
                       ba, ctr, rd);                                                          \
  else if (code == 1)                                                                         \
    hipLaunchKernelGGL((batch_kernel<SS_PREC_MIXED, T, 0, 0, 1, FS_PLAIN>), g, block, 0, st,  \
                       ba, ctr, rd);                                                          \
  else if (code == 2)                                                                         \
    hipLaunchKernelGGL((batch_kernel<SS_PREC_MIXED, T, 0, 0, 2, FS_PLAIN>), g, block, 0, st,  \
                       ba, ctr, rd);                                                          \
  else if (code == 3)                                                                         \
    hipLaunchKernelGGL((batch_kernel<SS_PREC_MIXED, T, 0, 0, 3, FS_PLAIN>), g, block, 0, st,  \
                       ba, ctr, rd);                                                          \
  else                                                                                        \
    hipLaunchKernelGGL((batch_kernel<SS_PREC_MIXED, T, 0, 0, 4, FS_PLAIN>), g, block, 0, st,  \
                       ba, ctr, rd)
#define SS_BATCH_C(P, T)                                                                      \
  if (P == SS_PREC_MIXED && cls == 0) {                                                       \
    SS_BATCH_PLAIN(T);                                                                        \
  } else if (code == 2) {                                                                     \
    hipLaunchKernelGGL((batch_kernel<P, T, -1, -1, 2>), g, block, 0, st, ba, ctr, rd);        \
  } else if (code == 3) {                                                                     \
    hipLaunchKernelGGL((batch_kernel<P, T, -1, -1, 3>), g, block, 0, st, ba, ctr, rd);        \
  } else if (code == 4) {                                                                     \
    hipLaunchKernelGGL((batch_kernel<P, T, -1, -1, 4>), g, block, 0, st, ba, ctr, rd);        \
  } else if (cls == 0) { SS_BATCH_K(P, T, 0, 0); }                                            \
  else if (cls == 1) { SS_BATCH_K(P, T, 1, 5); }                                              \
  else { SS_BATCH_K(P, T, -1, -1); }
#define SS_LAUNCH_BATCH(T)                                                                    \
  if (fp64) {                                                                                 \
    SS_BATCH_C(SS_PREC_FP64, T)                                                               \
  } else {                                                                                    \
    SS_BATCH_C(SS_PREC_MIXED, T)                                                              \
    if (rblocks[code] > 0) {                                                                  \
      if (code == 1)                                                                          \
        hipLaunchKernelGGL((batch_redo_kernel<T, 1>), rg, block, 0, st, ba, rd);              \
      else if (code == 2)                                                                     \
        hipLaunchKernelGGL((batch_redo_kernel<T, 2>), rg, block, 0, st, ba, rd);              \
      else if (code == 3)                                                                     \
        hipLaunchKernelGGL((batch_redo_kernel<T, 3>), rg, block, 0, st, ba, rd);              \
      else if (code == 4)                                                                     \
        hipLaunchKernelGGL((batch_redo_kernel<T, 4>), rg, block, 0, st, ba, rd);              \
    }                                                                                         \
  }
    if (tt == 4) {
      SS_LAUNCH_BATCH(4);
    } else if (tt == 8) {
      SS_LAUNCH_BATCH(8);
    } else {
      SS_LAUNCH_BATCH(16);
    }
#undef SS_LAUNCH_BATCH
#undef SS_BATCH_C
#undef SS_BATCH_PLAIN
#undef SS_BATCH_K
  }
  if (fp64 && fold)
    hipLaunchKernelGGL(batch_fold_kernel, dim3(cus > 4 ? cus / 4 : 1, 16), dim3(256), 0, st, ba);
}

static unsigned redo_grid(long items) {
  const long chunks = (items + kBlock - 1) / kBlock;
  return (unsigned)(chunks < 1024 ? chunks : 1024);
}

int64_t ss_batch_workspace_bytes(int precision, const ss_job* jobs_host, int n_jobs,
                                 int trace_tile) {
  if (n_jobs <= 0 || !jobs_host) return kWsHeader;
  if (trace_tile != 4 && trace_tile != 8 && trace_tile != 16) return -1;
  if (precision != SS_PREC_MIXED) return kWsHeader;
  return redo_bytes(batch_items(jobs_host, n_jobs, trace_tile), trace_tile);
}

int ss_batch_field(int precision, const ss_job* jobs_host, int n_jobs, void* jobs_device,
                   void* workspace, int64_t workspace_bytes, int trace_tile, int n_blocks,
                   void* stream) {
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (n_jobs < 0 || n_jobs > kMaxJobs) return fail(SS_ERANGE, "n_jobs out of range [0, 1024]");
  if (trace_tile != 4 && trace_tile != 8 && trace_tile != 16)
    return fail(SS_EINVAL, "trace_tile must be 4, 8 or 16");
  if (n_jobs == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!jobs_host || !jobs_device || !workspace) return fail(SS_EINVAL, "null pointer");
  if ((reinterpret_cast<uintptr_t>(workspace) & 7) != 0)
    return fail(SS_EINVAL, "workspace must be 8-byte aligned");
  for (int i = 0; i < n_jobs; ++i) {
    int r = validate_job(jobs_host[i], i);
    if (r) return r;
  }
  if (workspace_bytes < ss_batch_workspace_bytes(precision, jobs_host, n_jobs, trace_tile))
    return fail(SS_EINVAL, "workspace smaller than ss_batch_workspace_bytes()");
  long items[kJobCodes] = {};
  bool any_surface = false;
  for (int i = 0; i < n_jobs; ++i) {
    items[job_code(jobs_host[i])] += job_work(jobs_host[i], trace_tile);
    any_surface |= jobs_host[i].kind == 0;
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(jobs_device, jobs_host, sizeof(ss_job) * (size_t)n_jobs,
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail(SS_EHIP, std::string("job upload: ") + hipGetErrorString(e));
  const bool fp64 = precision == SS_PREC_FP64;
  const int nb = default_blocks(n_blocks);
  unsigned grid[kJobCodes], rblocks[kJobCodes] = {};
  for (int k = 0; k < kJobCodes; ++k) {
    grid[k] = (unsigned)(items[k] < nb ? items[k] : nb);
    if (!fp64 && k > 0) rblocks[k] = redo_grid(items[k]);
  }
  // the redo flags and masks hold the refracted classes' work ids in turn
  const Redo rd = fp64 ? Redo{} : redo_view(workspace, items[1] + items[2] + items[3] + items[4]);
  const BatchArgs ba{(const ss_job*)jobs_device, nullptr, n_jobs};
  launch_batch(fp64, ba, workspace, rd, trace_tile, grid, rblocks,
               batch_class(jobs_host, n_jobs), any_surface, st);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("batch launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int64_t ss_batch_table_bytes(int max_jobs) {
  if (max_jobs < 1 || max_jobs > kMaxJobs) return -1;
  return SS_BATCH_TABLE_HEADER + (int64_t)sizeof(ss_job) * max_jobs;
}

int ss_batch_pack(const ss_job* jobs_host, int n_jobs, int trace_tile, void* table_host,
                  int64_t table_bytes, int64_t* items, int* feature_class) {
  if (n_jobs < 0 || n_jobs > kMaxJobs) return fail(SS_ERANGE, "n_jobs out of range [0, 1024]");
  if (trace_tile != 4 && trace_tile != 8 && trace_tile != 16)
    return fail(SS_EINVAL, "trace_tile must be 4, 8 or 16");
  if (!table_host || (n_jobs > 0 && !jobs_host)) return fail(SS_EINVAL, "null pointer");
  if (table_bytes < ss_batch_table_bytes(n_jobs > 0 ? n_jobs : 1))
    return fail(SS_EINVAL, "table smaller than ss_batch_table_bytes(n_jobs)");
  for (int i = 0; i < n_jobs; ++i) {
    int r = validate_job(jobs_host[i], i);
    if (r) return r;
  }
  char* t = static_cast<char*>(table_host);
  std::memset(t, 0, SS_BATCH_TABLE_HEADER);
  std::memcpy(t, &n_jobs, sizeof(int));
  if (n_jobs > 0) std::memcpy(t + SS_BATCH_TABLE_HEADER, jobs_host, sizeof(ss_job) * n_jobs);
  long refr = 0, all = 0;
  for (int i = 0; i < n_jobs; ++i) {
    const long w = job_work(jobs_host[i], trace_tile);
    all += w;
    if (jobs_host[i].kind == 1) refr += w;
  }
  if (items) {
    items[0] = all - refr;
    items[1] = refr;
  }
  if (feature_class) *feature_class = batch_class(jobs_host, n_jobs);
  g_last_error.clear();
  return SS_OK;
}

int64_t ss_batch_step_workspace_bytes(int precision, int64_t max_items, int trace_tile) {
  if (trace_tile != 4 && trace_tile != 8 && trace_tile != 16) return -1;
  if (max_items < 0) return -1;
  if (precision != SS_PREC_MIXED) return kWsHeader;
  return redo_bytes(max_items, trace_tile);
}

int ss_batch_step(int precision, const void* table_device, void* workspace,
                  int64_t workspace_bytes, int64_t max_items, int trace_tile, int n_blocks,
                  int feature_class, void* stream) {
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (trace_tile != 4 && trace_tile != 8 && trace_tile != 16)
    return fail(SS_EINVAL, "trace_tile must be 4, 8 or 16");
  if (feature_class < 0 || feature_class > 2)
    return fail(SS_EINVAL, "feature_class must be 0, 1 or 2");
  if (!table_device || !workspace) return fail(SS_EINVAL, "null pointer");
  if ((reinterpret_cast<uintptr_t>(workspace) & 7) != 0 ||
      (reinterpret_cast<uintptr_t>(table_device) & 7) != 0)
    return fail(SS_EINVAL, "table and workspace must be 8-byte aligned");
  if (max_items < 1) return fail(SS_EINVAL, "max_items must be >= 1");
  if (workspace_bytes < ss_batch_step_workspace_bytes(precision, max_items, trace_tile))
    return fail(SS_EINVAL, "workspace smaller than ss_batch_step_workspace_bytes()");
  const bool fp64 = precision == SS_PREC_FP64;
  const int nb = default_blocks(n_blocks);
  unsigned grid[kJobCodes], rblocks[kJobCodes] = {};
  for (int k = 0; k < kJobCodes; ++k) grid[k] = (unsigned)nb;
  if (!fp64)
    for (int k = 1; k < kJobCodes; ++k) rblocks[k] = redo_grid(max_items);
  const Redo rd = fp64 ? Redo{} : redo_view(workspace, max_items);
  const char* t = static_cast<const char*>(table_device);
  const BatchArgs ba{reinterpret_cast<const ss_job*>(t + SS_BATCH_TABLE_HEADER),
                     reinterpret_cast<const int*>(t), 0};
  launch_batch(fp64, ba, workspace, rd, trace_tile, grid, rblocks, feature_class, true,
               (hipStream_t)stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("batch step launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_batch_overflowed(const void* workspace, int* overflowed, void* stream) {
  if (!workspace || !overflowed) return fail(SS_EINVAL, "null pointer");
  if ((reinterpret_cast<uintptr_t>(workspace) & 7) != 0)
    return fail(SS_EINVAL, "workspace must be 8-byte aligned");
  unsigned word = 0;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(&word, static_cast<const char*>(workspace) + kWsOverflow,
                                sizeof(word), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(SS_EHIP, std::string("overflow read: ") + hipGetErrorString(e));
  *overflowed = word != 0u;
  g_last_error.clear();
  return SS_OK;
}

int64_t ss_facets_workspace_bytes(int64_t ny, int64_t nx) {
  if (ny < 1 || nx < 1) return -1;
  const int64_t blocks = (ny * nx + kBlock - 1) / kBlock;
  return blocks * kStatCount * (int64_t)sizeof(double);
}

// lattice vertices per side: the column pass stages 2 rows + 32 taps in
// 64 KiB of LDS, the row pass 2 cols values
constexpr int64_t kGaussMaxSide = 4080;

int64_t ss_gaussian_heights_workspace_bytes(int64_t rows, int64_t cols) {
  if (rows < 1 || cols < 1) return -1;
  return rows * cols * (int64_t)sizeof(double);
}

int ss_gaussian_heights(uint64_t key_base, int64_t rows, int64_t cols, double spacing,
                        double corr_len, double sigma, double mean, double* heights,
                        void* workspace, int64_t workspace_bytes, void* stream) {
  if (rows < 1 || cols < 1) return fail(SS_EINVAL, "need at least one vertex per side");
  if (rows > kGaussMaxSide || cols > kGaussMaxSide)
    return fail(SS_ERANGE, "gaussian heights: at most 4080 vertices per side");
  if (!(spacing > 0.0) || !(corr_len > 0.0))
    return fail(SS_EINVAL, "spacing and corr_len must be > 0");
  if (!std::isfinite(sigma) || !std::isfinite(mean))
    return fail(SS_EINVAL, "sigma and mean must be finite");
  if (!heights || !workspace) return fail(SS_EINVAL, "null pointer");
  if (workspace_bytes < ss_gaussian_heights_workspace_bytes(rows, cols))
    return fail(SS_EINVAL, "workspace smaller than ss_gaussian_heights_workspace_bytes()");
  if ((reinterpret_cast<uintptr_t>(workspace) & 7) != 0)
    return fail(SS_EINVAL, "workspace must be 8-byte aligned");
  // normalisation 1 / sqrt(sum g^2) of the 2-D kernel = product of the
  // separable 1-D sums (minimal-image distances, as the passes' taps)
  const double a = corr_len / std::sqrt(2.0), a2 = a * a;
  auto sum_sq = [&](int64_t n) {
    double s = 0.0;
    for (int64_t d0 = 0; d0 < n; ++d0) {
      const int64_t d = d0 < n - d0 ? d0 : n - d0;
      const double c = (double)d * spacing;
      const double g = std::exp(-(c * c) / a2);
      s += g * g;
    }
    return s;
  };
  const double scale = sigma / std::sqrt(sum_sq(rows) * sum_sq(cols));
  hipStream_t st = (hipStream_t)stream;
  double* tmp = static_cast<double*>(workspace);
  hipLaunchKernelGGL(gauss_rows_kernel, dim3((unsigned)rows), dim3(kBlock),
                     (size_t)(2 * cols) * sizeof(double), st, (unsigned long long)key_base,
                     (long)cols, spacing, a2, tmp);
  const dim3 grid((unsigned)((cols + kBlock - 1) / kBlock),
                  (unsigned)((rows + kGaussRows - 1) / kGaussRows));
  hipLaunchKernelGGL(gauss_cols_kernel, grid, dim3(kBlock),
                     (size_t)(2 * rows + kGaussRows) * sizeof(double), st, tmp, (long)rows,
                     (long)cols, spacing, a2, scale, mean, heights);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SS_EHIP, std::string("gaussian heights launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_keyed_phasors(uint64_t key_base, int64_t ny, int64_t nx, int64_t key_row0,
                     int64_t key_col0, double* facets, int64_t facet_stride, void* stream) {
  if (ny < 1 || nx < 1) return fail(SS_EINVAL, "need at least one cell (ny, nx >= 1)");
  if (ny > (1L << 20) || nx > (1L << 20) || ny * nx > (1L << 31))
    return fail(SS_ERANGE, "lattice too large (ny * nx <= 2^31)");
  if (!facets) return fail(SS_EINVAL, "null pointer");
  if (facet_stride < ny * nx) return fail(SS_EINVAL, "facet_stride smaller than ny * nx");
  const long blocks = (long)((ny * nx + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(keyed_phasor_kernel, dim3((unsigned)blocks), dim3(kBlock), 0,
                     (hipStream_t)stream, (unsigned long long)key_base, (long)ny, (long)nx,
                     (long)key_row0, (long)key_col0, facets, (long)facet_stride);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SS_EHIP, std::string("keyed phasors launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_facets_from_heights(const double* heights, int64_t ny, int64_t nx, double spacing,
                           double x0, double y0, double* facets, int64_t facet_stride,
                           double* table, double* stats, void* workspace,
                           int64_t workspace_bytes, void* stream) {
  if (ny < 1 || nx < 1) return fail(SS_EINVAL, "need at least one cell (ny, nx >= 1)");
  if (ny > (1L << 20) || nx > (1L << 20) || ny * nx > (1L << 31))
    return fail(SS_ERANGE, "lattice too large (ny * nx <= 2^31)");
  if (!(spacing > 0.0)) return fail(SS_EINVAL, "spacing must be > 0");
  if (!heights || !facets || !stats || !workspace) return fail(SS_EINVAL, "null pointer");
  if (facet_stride < ny * nx) return fail(SS_EINVAL, "facet_stride smaller than ny * nx");
  if (workspace_bytes < ss_facets_workspace_bytes(ny, nx))
    return fail(SS_EINVAL, "workspace smaller than ss_facets_workspace_bytes()");
  if ((reinterpret_cast<uintptr_t>(workspace) & 7) != 0)
    return fail(SS_EINVAL, "workspace must be 8-byte aligned");
  const long blocks = (long)((ny * nx + kBlock - 1) / kBlock);
  hipStream_t st = (hipStream_t)stream;
  double* partials = static_cast<double*>(workspace);
  hipLaunchKernelGGL(facets_from_heights_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st,
                     heights, (long)ny, (long)nx, spacing, x0, y0, facets, (long)facet_stride,
                     table, partials);
  hipLaunchKernelGGL(facet_stats_kernel, dim3(1), dim3(kBlock), 0, st, partials, blocks, stats);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("facet build launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_trace_diagnostics(const double* facets, int64_t facet_stride, int64_t n_facets,
                         const int64_t* facet_index, const double* traces,
                         int64_t trace_stride, int64_t n_traces, double t0, double dt,
                         int64_t n_samples, double c, int64_t* nadir_index, double* nadir_z,
                         int64_t* first_index, double* first_twtt, double* first_xyz,
                         void* stream) {
  if (n_facets < 0 || n_traces < 0) return fail(SS_EINVAL, "negative sizes");
  if (facet_stride < n_facets || trace_stride < n_traces)
    return fail(SS_EINVAL, "stride smaller than count");
  if (!(dt > 0.0) || !(c > 0.0) || n_samples <= 0)
    return fail(SS_EINVAL, "dt, c and n_samples must be > 0");
  if (n_traces == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!traces || !nadir_index || !nadir_z || !first_index || !first_twtt || !first_xyz ||
      (n_facets > 0 && !facets))
    return fail(SS_EINVAL, "null pointer");
  const long blocks = (long)((n_traces + kDiagTT - 1) / kDiagTT);
  hipLaunchKernelGGL(trace_diag_kernel, dim3((unsigned)blocks), dim3(kBlock), 0,
                     (hipStream_t)stream, facets, (long)facet_stride, (long)n_facets,
                     reinterpret_cast<const long*>(facet_index), traces, (long)trace_stride,
                     (long)n_traces, t0, t0 + (double)n_samples * dt, c,
                     reinterpret_cast<long*>(nadir_index), nadir_z,
                     reinterpret_cast<long*>(first_index), first_twtt, first_xyz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("diagnostics launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int64_t ss_compress_fft_size(int64_t n_samples, int64_t m) {
  if (n_samples <= 0 || m < 0) return 0;
  return compress_fft_size((long)n_samples, (long)m);
}

int ss_pulse_spectrum(const double* pulse, int64_t m, int64_t nfft, double* spectrum,
                      void* stream) {
  if (!pulse || !spectrum) return fail(SS_EINVAL, "null pointer");
  if (m < 0) return fail(SS_EINVAL, "negative m");
  if (nfft < 256 || nfft > kFftMax || (nfft & (nfft - 1)) != 0)
    return fail(SS_ERANGE, "nfft must be a power of two in [256, 8192]");
  if (2 * m + 1 > nfft) return fail(SS_ERANGE, "pulse longer than nfft");
  const int logn = (int)std::lround(std::log2((double)nfft));
  hipLaunchKernelGGL(pulse_spectrum_kernel, dim3(1), dim3(kFftThreads), 0, (hipStream_t)stream,
                     (const cd*)pulse, (long)(2 * m + 1), (int)nfft, logn, (cd*)spectrum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("spectrum launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_range_compress_fft(const double* in, int64_t n_rows, int64_t n_samples,
                          const double* spectrum, int64_t m, int64_t nfft, double* out,
                          void* stream) {
  if (n_rows < 0 || n_samples < 0 || m < 0) return fail(SS_EINVAL, "negative sizes");
  if (n_rows == 0 || n_samples == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!in || !out || !spectrum) return fail(SS_EINVAL, "null pointer");
  if (in == out) return fail(SS_EINVAL, "in and out must not alias");
  if (nfft < 256 || nfft > kFftMax || (nfft & (nfft - 1)) != 0)
    return fail(SS_ERANGE, "nfft must be a power of two in [256, 8192]");
  const long L = nfft - (2 * m + 1) + 1;
  if (L < 1) return fail(SS_ERANGE, "pulse longer than nfft");
  if (n_rows > 65535) return fail(SS_ERANGE, "n_rows > 65535: split the call");
  const long blocks = (n_samples + L - 1) / L;
  const int logn = (int)std::lround(std::log2((double)nfft));
  const dim3 grid((unsigned)blocks, (unsigned)n_rows);
  hipStream_t st = (hipStream_t)stream;
#define SS_FFT(CAP)                                                                            \
  hipLaunchKernelGGL((compress_fft_kernel<CAP>), grid, dim3(kFftThreads), 0, st, (const cd*)in, \
                     (long)n_samples, (const cd*)spectrum, (long)m, (int)nfft, logn, (cd*)out)
  if (nfft <= 2048) {
    SS_FFT(2048);
  } else if (nfft <= 4096) {
    SS_FFT(4096);
  } else {
    SS_FFT(kFftMax);
  }
#undef SS_FFT
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("compress launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_range_compress(const double* in, int64_t n_rows, int64_t n_samples, const double* pulse,
                      int64_t m, double* out, void* stream) {
  if (n_rows < 0 || n_samples < 0 || m < 0) return fail(SS_EINVAL, "negative sizes");
  if (n_rows == 0 || n_samples == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!in || !out || !pulse) return fail(SS_EINVAL, "null pointer");
  if (in == out) return fail(SS_EINVAL, "in and out must not alias");
  if (n_rows > 65535) return fail(SS_ERANGE, "n_rows > 65535: split the call");
  hipStream_t st = (hipStream_t)stream;
  dim3 grid((unsigned)((n_samples + kBlock - 1) / kBlock), (unsigned)n_rows);
  hipLaunchKernelGGL(compress_kernel, grid, dim3(kBlock), 0, st, (const double2*)in,
                     (long)n_samples, (const double2*)pulse, (long)m, (double2*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("compress launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_slow_time_boxcar(const double* in, int64_t rows, int64_t cols, int64_t window,
                        int64_t stride, int64_t n_out, double scale, double* out,
                        void* stream) {
  if (rows < 0 || cols < 0 || n_out < 0) return fail(SS_EINVAL, "negative sizes");
  if (window < 1 || stride < 1) return fail(SS_EINVAL, "window and stride must be >= 1");
  if (n_out > 0 && (n_out - 1) * stride + window > rows)
    return fail(SS_EINVAL, "the last window runs past the input rows");
  if (n_out == 0 || cols == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!in || !out) return fail(SS_EINVAL, "null pointer");
  if (in == out) return fail(SS_EINVAL, "in and out must not alias");
  const long gy = n_out < 65535 ? n_out : 65535;
  dim3 grid((unsigned)((cols + kBlock - 1) / kBlock), (unsigned)gy);
  hipLaunchKernelGGL(boxcar_kernel, grid, dim3(kBlock), 0, (hipStream_t)stream, in, (long)cols,
                     (long)window, (long)stride, (long)n_out, scale, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("boxcar launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

int ss_focused_sar(int precision, const double* field, int64_t n_traces, int64_t n_bins,
                   const double* twtt, const double* s, double c, double k, double aperture_m,
                   int window, double* out, void* stream) {
  if (precision != SS_PREC_FP64 && precision != SS_PREC_MIXED)
    return fail(SS_EINVAL, "precision must be SS_PREC_FP64 or SS_PREC_MIXED");
  if (n_traces < 0 || n_bins < 0) return fail(SS_EINVAL, "negative sizes");
  if (n_traces > 0 && n_bins < 2) return fail(SS_EINVAL, "need at least two bins");
  if (window != 0 && window != 1) return fail(SS_EINVAL, "window must be 0 (none) or 1 (hann)");
  if (!(aperture_m > 0.0) || !(c > 0.0)) return fail(SS_EINVAL, "aperture_m and c must be > 0");
  if (n_traces == 0) {
    g_last_error.clear();
    return SS_OK;
  }
  if (!field || !twtt || !s || !out) return fail(SS_EINVAL, "null pointer");
  if (field == out) return fail(SS_EINVAL, "field and out must not alias");
  const long gy = n_traces < 65535 ? n_traces : 65535;
  dim3 grid((unsigned)((n_bins + kBlock - 1) / kBlock), (unsigned)gy);
  hipStream_t st = (hipStream_t)stream;
  if (precision == SS_PREC_FP64)
    hipLaunchKernelGGL(focus_kernel<SS_PREC_FP64>, grid, dim3(kBlock), 0, st, field,
                       (long)n_traces, (long)n_bins, twtt, s, c, k, aperture_m / 2.0, window,
                       out);
  else
    hipLaunchKernelGGL(focus_kernel<SS_PREC_MIXED>, grid, dim3(kBlock), 0, st, field,
                       (long)n_traces, (long)n_bins, twtt, s, c, k, aperture_m / 2.0, window,
                       out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SS_EHIP, std::string("focus launch: ") + hipGetErrorString(e));
  g_last_error.clear();
  return SS_OK;
}

}  // extern "C"

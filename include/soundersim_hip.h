/*
 * soundersim_hip.h -- C ABI of the MI355X (gfx950) coherent facet kernel.
 *
 * Plain extern "C" functions over device pointers and sizes: no C++ types,
 * no torch types, no exceptions cross this boundary.  Every entry point
 * returns SS_OK (0) or a negative SS_E* code; ss_last_error() then gives a
 * message (thread-local).  All work is enqueued on the caller's HIP stream
 * (``stream`` is a hipStream_t, NULL = default stream); the caller owns every
 * buffer and the library never frees caller memory.
 *
 * The reference (soundersim, englacial/radar-sounder-simulator) has no
 * native layer: its path sits behind plain Python functions.  Each entry
 * point below replaces one of them and is bound from Python by ctypes
 * (radar_sounder_simulator_amd/_lib.py) behind a shim with the reference's
 * signature.
 *
 * Data layout (device memory, float64 structure-of-arrays):
 *   facets  [SS_FROW_COUNT][facet_stride]  rows SS_FROW_* below; the
 *           optional rows are read only when the matching flag is set.
 *           Facets are expected in a spatially compact order (the Python
 *           shim Morton-sorts them); the kernels take 256 consecutive
 *           facets as one tile.
 *   traces  [SS_TROW_COUNT][trace_stride]  rows SS_TROW_* below.
 *   field   (n_traces, n_seg) complex128 interleaved (re, im), n_seg =
 *           n_samples (x2 with SS_SPLIT_SIDES: [left | right] halves);
 *           incoherent refracted output uses the same layout with the
 *           power in the real part and a zero imaginary part.
 *   dropped (n_traces,) float64.
 *   pair_bins (optional, may be NULL): (n_traces, n_facets) int32, the
 *           floor bin of every pair before the side offset, or
 *           SS_BIN_INVALID for a refracted pair whose path is invalid.
 * Outputs are zeroed by the call (no memset node: a fill kernel).
 */
#ifndef SOUNDERSIM_HIP_H
#define SOUNDERSIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_OK 0
#define SS_EINVAL (-1)
#define SS_EHIP (-2)
#define SS_ERANGE (-3)

/* precision */
#define SS_PREC_FP64 0  /* plain float64 everywhere (follows the oracle)    */
#define SS_PREC_MIXED 1 /* float64 range/bin/phase cycles + Snell polish,
                           float32 amplitudes, sinc, antenna, roughness    */

/* flags */
#define SS_SPLIT_SIDES 0x001
#define SS_INTERP_BINS 0x002
#define SS_TAPER 0x004
#define SS_D_PHI_AREA 0x008
#define SS_ROUGH 0x010
#define SS_GAMMA_FACET 0x020
#define SS_COHERENT 0x040     /* refracted: coherent (field) vs incoherent  */
#define SS_CROSS_ROUGH 0x080  /* refracted: crossing roughness sig_c        */
#define SS_DIFFUSE 0x100      /* refracted: diffuse channel                 */
#define SS_REDO_ALL 0x200     /* refracted, testing: every mixed-tier pair  */
                              /* takes the float64 redo launch              */

/* antenna pattern kinds (antenna.py gain_fn) */
#define SS_PAT_ISOTROPIC 0
#define SS_PAT_DIPOLE 1
#define SS_PAT_ARRAY 2
#define SS_PAT_ARRAY_TAPERED 3
#define SS_PAT_FINITE_DIPOLE 4
#define SS_PAT_TABULATED 5

/* roughness correlation functions */
#define SS_ACF_GAUSSIAN 0
#define SS_ACF_EXPONENTIAL 1

/* facet rows */
#define SS_FROW_CX 0
#define SS_FROW_NX 3
#define SS_FROW_AREA 6
#define SS_FROW_E1X 7
#define SS_FROW_E2X 10
#define SS_FROW_GAMMA 13
#define SS_FROW_PH_RE 14
#define SS_FROW_PH_IM 15
#define SS_FROW_DIF_AMP 16
#define SS_FROW_DIF_RE 17
#define SS_FROW_DIF_IM 18
#define SS_FROW_COUNT 19

/* trace rows */
#define SS_TROW_PX 0
#define SS_TROW_UX 3      /* cross-track unit vector (split sides)        */
#define SS_TROW_RREF 6    /* surface: |p - mean(centres)| (phase fold-back) */
#define SS_TROW_PV0 7     /* pattern vector (axis / boresight)           */
#define SS_TROW_PV1 10    /* second pattern vector (array boresight)     */
#define SS_TROW_COUNT 13

/* crossed-interface descriptor rows (refracted), one row of 16 per iface */
#define SS_IDESC_NY 0
#define SS_IDESC_NX 1
#define SS_IDESC_COEF 2   /* (3, 2) row-major: coef[0..5]                */
#define SS_IDESC_MP 8
#define SS_IDESC_MN 11
#define SS_IDESC_SIGC 14
#define SS_IDESC_TABOFF 15 /* first row of this interface in ``tables``   */
#define SS_IDESC_COUNT 16

#define SS_BIN_INVALID (-2147483647 - 1)
#define SS_MAX_CROSSED 64
#define SS_MAX_JOINT 32
#define SS_MAX_TERMS 300
#define SS_MAX_PATTERN_TABLE 4096

/* Kernel-wide constants table (device, float64), see ss_fill_constants. */
#define SS_CONST_WEIDEMAN_L 0
#define SS_CONST_WEIDEMAN_A 1   /* 32 coefficients, highest power first    */
#define SS_CONST_LGAMMA 33      /* lgamma(m + 1), m = 0 .. SS_MAX_TERMS     */
#define SS_CONST_COUNT (33 + SS_MAX_TERMS + 1)

typedef struct ss_surface_params {
  double k;         /* carrier wavenumber 2 pi / lambda (rad/m)             */
  double gamma;     /* scalar field reflection coefficient                 */
  double t0;        /* window start (s)                                    */
  double dt;        /* sample spacing (s)                                  */
  double c;         /* propagation speed (m/s)                             */
  double taper_s;   /* off-specular taper s_eff (SS_TAPER)                 */
  double sigma;     /* roughness rms height (m) (SS_ROUGH)                 */
  double corr_len;  /* roughness correlation length (m)                    */
  double pat_a;     /* array: n_elements; finite dipole: length_lam        */
  double pat_b[4];  /* array: spacing_lam, field scale, q, element enabled */
  int32_t n_samples;
  int32_t flags;
  int32_t n_terms;  /* D_Phi series length (SS_ROUGH)                      */
  int32_t acf;      /* SS_ACF_*                                            */
  int32_t pattern;  /* SS_PAT_*                                            */
  int32_t pat_n;    /* tabulated samples / array_tapered elements          */
} ss_surface_params;

typedef struct ss_refracted_params {
  double k0;        /* vacuum wavenumber (coherent)                        */
  double gamma;     /* scalar target reflection coefficient               */
  double t0, dt, c;
  double taper_s;
  double sigma;     /* target roughness rms height                         */
  double corr_len;
  double n_exp;     /* diffuse cos^n exponent                              */
  double pat_a;
  double pat_b[4];
  int32_t n_samples;
  int32_t flags;
  int32_t n_crossed;  /* interfaces above the target (1..SS_MAX_CROSSED)   */
  int32_t n_terms;
  int32_t acf;
  int32_t pattern;
  int32_t pat_n;
  int32_t refraction;       /* 0 sequential chain, 1 joint solve, 2 joint
                               solve over a flat stack (see below)       */
  int32_t k_pad;            /* joint: index-matched pads at the top       */
  int32_t joint_newton;     /* joint: damped Newton steps (reference 6)   */
  int32_t joint_backtrack;  /* joint: step halvings (reference 4)         */
  int32_t reserved;
} ss_refracted_params;

/* Incoherent (power-summing) surface kernel parameters. */
typedef struct ss_incoherent_params {
  double t0, dt, c;
  double pat_a;     /* antenna parameters as in ss_surface_params          */
  double pat_b[4];
  int32_t n_samples;
  int32_t flags;    /* SS_SPLIT_SIDES only                                 */
  int32_t pattern;  /* SS_PAT_*                                            */
  int32_t pat_n;
} ss_incoherent_params;

/* One layer of one request in a batched launch (ss_batch_field).  kind 0:
 * surface layer (uses ``surface``), kind 1: refracted layer (uses
 * ``refracted`` and the interface pointers) with the solver that
 * refracted.refraction names: 0 sequential chain, 1 joint solve (stack
 * incl. pads at most SS_MAX_JOINT deep), 2 joint solve over a flat stack
 * (as ss_refracted_field).  All pointers are device pointers with the
 * layouts above; field and dropped must be ordinary (coarse-grained) device
 * allocations such as hipMalloc or torch give: the kernels accumulate into
 * them with hardware float64 atomics.  Outputs are zeroed by the call. */
typedef struct ss_job {
  int32_t kind;
  int32_t precision_unused;   /* the batch precision is the call's          */
  int64_t facet_stride, n_facets;
  int64_t trace_stride, n_traces;
  const double* facets;
  const double* traces;
  double* field;
  double* dropped;
  const double* iface_desc;
  const double* tables;
  const double* legs;
  const double* pattern_table;
  const double* constants;
  ss_surface_params surface;
  ss_refracted_params refracted;
} ss_job;

/* Library version string. */
const char* ss_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* ss_last_error(void);

/* Number of HIP devices visible (0 when none). */
int ss_device_count(void);

/* Fill the kernel constants table (SS_CONST_COUNT float64, host memory):
 * Weideman (1994) N = 32 rational-approximation constants for the Faddeeva
 * function and lgamma(m + 1).  Replaces roughness.py:55-64 (_weideman_coeffs)
 * and the gammaln call of roughness.py:172.  Host-only, no device work. */
int ss_fill_constants(double* host_out);

/* Surface-layer coherent field.  Replaces kernels/coherent.py:247-362
 * (coherent_cluttergram) and its device program _coherent_fn (:171-244).
 * pattern_table: tabulated [theta (pat_n) | gain (pat_n)] or array_tapered
 * [tx (pat_n) | rx (pat_n)] float64 device array, else NULL.
 * constants: SS_CONST_COUNT float64 device array (required with SS_ROUGH). */
int ss_surface_field(int precision, const ss_surface_params* params,
                     const double* facets, int64_t facet_stride,
                     int64_t n_facets, const double* traces,
                     int64_t trace_stride, int64_t n_traces,
                     const double* pattern_table, const double* constants,
                     double* field, double* dropped, int32_t* pair_bins,
                     void* stream);

/* Incoherent surface power.  Replaces kernels/incoherent.py:66-105
 * (incoherent_cluttergram) and its device program _incoherent_fn (:31-64):
 * per pair (A cos)^2 / r^4, times g^4 for a non-isotropic antenna, binned at
 * floor((2r/c - t0)/dt); out-of-window power goes to ``dropped``.  Reads
 * the centre, normal and area rows of ``facets`` only.  power: (n_traces,
 * n_seg) float64 (real, not interleaved), n_seg = n_samples (x2 with
 * SS_SPLIT_SIDES: [left | right]); zeroed by the call.  pair_bins as for
 * ss_surface_field.  SS_PREC_MIXED: float64 range, bin and side, float32
 * power and gain (bins equal SS_PREC_FP64's). */
int ss_incoherent_power(int precision, const ss_incoherent_params* params,
                        const double* facets, int64_t facet_stride,
                        int64_t n_facets, const double* traces,
                        int64_t trace_stride, int64_t n_traces,
                        const double* pattern_table, double* power,
                        double* dropped, int32_t* pair_bins, void* stream);

/* Refracted-path field of one target interface under n_crossed interfaces,
 * sequential two-point chain or joint N-interface solve (refraction = 1,
 * stack incl. pads at most SS_MAX_JOINT deep).  refraction = 2 is the joint
 * solve for a stack whose real interfaces below the surface (indices
 * k_pad + 1 ..) are horizontal planes: every table cell of each has normal
 * (0, 0, +nz) and the same z, which the caller guarantees (the kernel reads
 * each plane's z from its first cell).  It gives the joint solve's
 * stationary path with two unknowns instead of 2 * n_crossed, needs no
 * per-interface storage and takes stacks up to SS_MAX_CROSSED deep.  Replaces
 * kernels/multilayer.py:594-770 (refracted_cluttergram) and _refracted_fn
 * (:333-591), refraction_joint.py:249-322 (joint_crossings).  iface_desc: (n_crossed, SS_IDESC_COUNT) float64 device;
 * tables: concatenated (sum ny*nx, 6) float64 [centre, normal] device;
 * legs: (3, n_crossed + 1) float64 device rows [n, eps, att_db_per_km].
 * workspace: caller-owned device memory, 8-byte aligned, of at least
 * ss_refracted_workspace_bytes(precision, n_facets, n_traces) bytes (may be
 * NULL when that is 0, i.e. for SS_PREC_FP64).  It holds the per-work-item
 * flags and lane masks of the pairs the mixed tier hands to its float64 redo
 * pass (rewritten by every call, nothing to clear), so concurrent calls need
 * distinct workspaces. */
int ss_refracted_field(int precision, const ss_refracted_params* params,
                       const double* facets, int64_t facet_stride,
                       int64_t n_facets, const double* traces,
                       int64_t trace_stride, int64_t n_traces,
                       const double* iface_desc, const double* tables,
                       const double* legs, const double* pattern_table,
                       const double* constants, double* out, double* dropped,
                       int32_t* pair_bins, void* workspace, int64_t workspace_bytes,
                       void* stream);

/* ss_refracted_field without a workspace argument, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * mixed tier's redo flags and lane masks go to a library-owned device array
 * of SS_NOALLOC_SCRATCH_BYTES per GPU (held from library load until the
 * process ends); its address is resolved once per device on the first call
 * made with that device current.  The call allocates nothing.  Inputs,
 * outputs and semantics are those of ss_refracted_field; a call whose
 * ss_refracted_workspace_bytes() exceeds SS_NOALLOC_SCRATCH_BYTES fails with
 * SS_ERANGE before any work is enqueued.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (different streams or threads
 * with no ordering between them) race; calls ordered on one stream do not. */
#define SS_NOALLOC_SCRATCH_BYTES (32LL << 20)
int ss_refracted_field_noalloc(int precision, const ss_refracted_params* params,
                               const double* facets, int64_t facet_stride, int64_t n_facets,
                               const double* traces, int64_t trace_stride, int64_t n_traces,
                               const double* iface_desc, const double* tables,
                               const double* legs, const double* pattern_table,
                               const double* constants, double* out, double* dropped,
                               int32_t* pair_bins, void* stream);

/* Workspace bytes ss_refracted_field needs: 0 for SS_PREC_FP64 or an empty
 * call; for SS_PREC_MIXED 256 + one flag byte per (16-trace, 256-facet) work
 * item (rounded up to 8) + 512 bytes of lane masks per work item.  No
 * counterpart in the reference (its kernels run one precision). */
int64_t ss_refracted_workspace_bytes(int precision, int64_t n_facets, int64_t n_traces);

/* Batched persistent launch over uneven jobs: every (job, trace tile, facet
 * tile) work item of ``n_jobs`` jobs is drained from one device work counter
 * by ``n_blocks`` persistent workgroups (0 = a default of 4 per CU).  The
 * job table is copied from host memory into ``jobs_device`` (caller-owned
 * device buffer of at least n_jobs * sizeof(ss_job) bytes).  ``workspace``
 * (caller-owned device memory, 8-byte aligned, at least
 * ss_batch_workspace_bytes() bytes) holds the work counter and, for
 * SS_PREC_MIXED, the float64 redo flags and lane masks; the counter is reset
 * on the stream, so the call allocates nothing.  trace_tile: traces per workgroup (4, 8 or 16;
 * chosen by the autotuner).  No counterpart in the reference (its requests
 * run one jit call after another); serves batched and continuous-batching
 * callers. */
int ss_batch_field(int precision, const ss_job* jobs_host, int n_jobs, void* jobs_device,
                   void* workspace, int64_t workspace_bytes, int trace_tile, int n_blocks,
                   void* stream);

/* Workspace bytes of ss_batch_field for these jobs (host table) and trace
 * tile: 256 for SS_PREC_FP64; for SS_PREC_MIXED plus one flag byte per work
 * item (rounded up to 8) and 32 * trace_tile bytes of lane masks per work
 * item; -1 for an invalid trace_tile.  No counterpart in the reference. */
int64_t ss_batch_workspace_bytes(int precision, const ss_job* jobs_host, int n_jobs,
                                 int trace_tile);

/* Device-resident batched step (capturable in a HIP graph).  A job table in
 * device memory -- SS_BATCH_TABLE_HEADER bytes holding the int32 job count,
 * then the ss_job records -- is read by the kernels themselves: outputs are
 * zeroed, each job class (surface, and the refracted jobs of each solver)
 * runs as a persistent launch, the float64 redo and the fold-back follow,
 * all with grids fixed at the call, so the launch sequence can be captured
 * once and replayed after each refresh of the table (e.g. a copy from a
 * pinned host table written by ss_batch_pack).  max_items bounds the
 * refracted work items (items[1] of ss_batch_pack) a table may hold: it
 * sizes the mixed tier's redo record.  A mixed step over a table with more
 * refracted items cannot defer their unresolved pairs: it sets the
 * workspace's overflow word (ss_batch_overflowed) and those pairs are
 * missing from the outputs, so callers check items[1] <= max_items before
 * the step (the request loop does) or read the word after it.
 * feature_class (from ss_batch_pack) must cover the jobs.  No counterpart
 * in the reference; serves the request loop. */
#define SS_BATCH_TABLE_HEADER 256

/* Bytes of a job table for up to max_jobs jobs (-1 if out of range). */
int64_t ss_batch_table_bytes(int max_jobs);

/* Validate n_jobs jobs and write the table (header + records) into
 * table_host (at least ss_batch_table_bytes(n_jobs) bytes; pinned memory
 * for captured replays).  items[0], items[1]: work items of the surface and
 * refracted jobs at this trace_tile; feature_class: 0 (no roughness,
 * isotropic, and every job plain: no split sides, interpolated bins,
 * diffuse term, per-facet gamma or forced redo, one crossing per
 * sequential job, joint stacks of at most three; the mixed launches fold
 * those flags at compile time), 1 (at most area-only D_Phi, tabulated
 * pattern) or 2.  A step graph captured for a class runs every lower one. */
int ss_batch_pack(const ss_job* jobs_host, int n_jobs, int trace_tile, void* table_host,
                  int64_t table_bytes, int64_t* items, int* feature_class);

/* Workspace bytes of ss_batch_step for up to max_items refracted work items. */
int64_t ss_batch_step_workspace_bytes(int precision, int64_t max_items, int trace_tile);

int ss_batch_step(int precision, const void* table_device, void* workspace,
                  int64_t workspace_bytes, int64_t max_items, int trace_tile, int n_blocks,
                  int feature_class, void* stream);

/* Whether the last batched launch on ``workspace`` (ss_batch_field or
 * ss_batch_step, mixed tier) met more refracted work items than its redo
 * record holds: *overflowed = 1 then (its result is incomplete), else 0.
 * Waits for ``stream``.  No counterpart in the reference. */
int ss_batch_overflowed(const void* workspace, int* overflowed, void* stream);

/* sizeof(ss_job) as compiled into the library (for sizing jobs_device and
 * checking a binding's struct mirror).  No counterpart in the reference. */
int64_t ss_sizeof_job(void);

/* Fast-time 'same' convolution of n_rows rows of n_samples complex128
 * samples with a (2m+1)-tap complex128 pulse whose zero lag is tap m,
 * direct summation in float64.  Replaces waveform.py:126-138
 * (convolve_fast_time) as used by apply_waveform (:141-160).  in/out may
 * not alias. */
int ss_range_compress(const double* in, int64_t n_rows, int64_t n_samples,
                      const double* pulse, int64_t m, double* out,
                      void* stream);

/* FFT size ss_range_compress_fft uses for rows of n_samples and a (2m+1)-tap
 * pulse: a power of two in [256, 8192], or 0 when the pulse is too long for
 * the LDS-resident transform (use ss_range_compress).  No counterpart in
 * the reference (numpy picks its FFT length itself, waveform.py:134). */
int64_t ss_compress_fft_size(int64_t n_samples, int64_t m);

/* Spectrum of a (2m+1)-tap complex128 pulse for ss_range_compress_fft:
 * nfft complex128 values (device), in the transform's internal
 * (bit-reversed) order, scaled by 1/nfft.  Computed once per pulse. */
int ss_pulse_spectrum(const double* pulse, int64_t m, int64_t nfft, double* spectrum,
                      void* stream);

/* The same 'same' convolution as ss_range_compress by FFT overlap-save in
 * float64 (the reference's method, waveform.py:126-138): each workgroup
 * transforms nfft input samples in LDS, multiplies by the pulse spectrum
 * and keeps the nfft - 2m wrap-free outputs.  in/out may not alias. */
int ss_range_compress_fft(const double* in, int64_t n_rows, int64_t n_samples,
                          const double* spectrum, int64_t m, int64_t nfft, double* out,
                          void* stream);

/* Facets of a regular vertex lattice built on the device.  heights:
 * (ny+1, nx+1) float64 device, row-major (rows along y, columns along x),
 * vertices at ((j - nx/2) * spacing + x0, (i - ny/2) * spacing + y0).  Cell
 * (r, c) becomes a facet by the reference's corner math (scene.py:158-176,
 * compare/brute_force_layered.py:44-71: mean edges e1 (along +column) and
 * e2 (along +row), centre = vertex mean, normal = e1 x e2 oriented +z,
 * area = |e1 x e2|) written to facet rows SS_FROW_CX..SS_FROW_E2X + 2 at
 * position ss order: 16x16-cell tiles, tiles row-major, cells row-major in
 * a tile (r' = r mod 16, c' = c mod 16, h, w the tile's rows and columns:
 * pos = (r/16)*16*nx + (c/16)*16*h + r'*w + c').  Rows from SS_FROW_GAMMA on
 * are left to the caller.  table (nullable): (ny*nx, 6) float64 device
 * [centre, normal] in grid order, the crossed-interface lookup table of
 * ss_refracted_field.  stats: 10 float64 device values [sum A, sum A*centre
 * (3), sum A*normal (3), sum centre (3)] (the area-weighted mean plane and
 * the centre mean), summed in a fixed order.  workspace: caller memory of
 * ss_facets_workspace_bytes(ny, nx) bytes, 8-byte aligned.  Replaces the
 * host-side tessellation the reference runs before each kernel call; no
 * counterpart as a device routine. */
int ss_facets_from_heights(const double* heights, int64_t ny, int64_t nx, double spacing,
                           double x0, double y0, double* facets, int64_t facet_stride,
                           double* table, double* stats, void* workspace,
                           int64_t workspace_bytes, void* stream);

/* Workspace bytes of ss_facets_from_heights (per-block partial sums), or -1
 * for an empty lattice. */
int64_t ss_facets_workspace_bytes(int64_t ny, int64_t nx);

/* Seeded Gaussian-ACF height lattice built on the device (replaces
 * compare/haynes.py:94-112, gaussian_surface, the reference's generator of
 * synthetic DEMs): heights (rows, cols) float64 device, row-major,
 *   heights = mean + sigma * (w (*) g) / sqrt(sum g^2),
 * the circular convolution of white noise w with g(rho) = exp(-rho^2/a^2),
 * a = corr_len / sqrt(2), on minimal-image distances at `spacing` (unit
 * variance, correlation exp(-rho^2/corr_len^2)).  The noise is keyed, not
 * drawn from a stream: w[i, j] = sqrt(-2 log u1) cos(2 pi u2), u1 and u2
 * from the splitmix64 keying of roughness.py:207-233 with key (i, j) and
 * stream base key_base (the host derives it from a seed with numpy
 * SeedSequence(seed).generate_state(1)).  The convolution runs as two
 * direct separable circular passes in float64 (the reference multiplies
 * FFTs; results agree to rounding).  workspace: caller memory of
 * ss_gaussian_heights_workspace_bytes(rows, cols) bytes, 8-byte aligned.
 * rows, cols <= 4080 (SS_ERANGE beyond). */
int ss_gaussian_heights(uint64_t key_base, int64_t rows, int64_t cols, double spacing,
                        double corr_len, double sigma, double mean, double* heights,
                        void* workspace, int64_t workspace_bytes, void* stream);

/* Workspace bytes of ss_gaussian_heights (the row pass's output), or -1. */
int64_t ss_gaussian_heights_workspace_bytes(int64_t rows, int64_t cols);

/* Keyed speckle phasors of an (ny, nx)-cell lattice written into facet rows
 * SS_FROW_PH_RE / SS_FROW_PH_IM at the ss_facets_from_heights positions
 * (replaces roughness.py:207-233, speckle_phasors_for_keys, for the keys
 * (key_row0 + r, key_col0 + c) of cell (r, c)): (e1 + i e2) / sqrt 2 by
 * splitmix64 keying and Box-Muller, rounded to complex64 as the reference
 * returns it, stored as float64.  key_base as in ss_gaussian_heights. */
int ss_keyed_phasors(uint64_t key_base, int64_t ny, int64_t nx, int64_t key_row0,
                     int64_t key_col0, double* facets, int64_t facet_stride, void* stream);

/* Per-trace diagnostics over one interface's facets, float64: the facet
 * whose centre is horizontally nearest the platform (nadir_index, its centre
 * z in nadir_z) and the earliest exact arrival 2|p - centre|/c inside
 * [t0, t0 + n_samples*dt) (first_index, first_twtt, its centre in first_xyz
 * (T, 3)); ties go to the lowest caller index, as numpy's argmin.
 * facet_index (nullable, device int64 (n_facets,)): the caller's index of
 * each facet position (tie-breaks and the returned indices), else the
 * position itself.  No facet in the window: first_index -1, first_twtt +inf,
 * first_xyz NaN (the reference then reports facet 0, output.py:79-82).
 * Replaces output.py:69-82 (first_return_*) and the nearest-facet step of
 * simulate.py:398-412 (_nadir_twtt_layers), both O(T F) host arrays in the
 * reference. */
int ss_trace_diagnostics(const double* facets, int64_t facet_stride, int64_t n_facets,
                         const int64_t* facet_index, const double* traces,
                         int64_t trace_stride, int64_t n_traces, double t0, double dt,
                         int64_t n_samples, double c, int64_t* nadir_index, double* nadir_z,
                         int64_t* first_index, double* first_twtt, double* first_xyz,
                         void* stream);

/* Along-track boxcar (processing.py:143-238: presum, unfocused_sar,
 * multilook): out[k, c] = scale * sum_{j < window} in[k * stride + j, c] for
 * k < n_out, on a row-major (rows, cols) float64 device array (a complex
 * interleaved (rows, n) array is cols = 2 n).  presum: stride = window, scale
 * 1 (coherent sum of the field); unfocused SAR: stride 1; multilook: stride =
 * window, scale 1 / window on power.  (n_out - 1) * stride + window <= rows. */
int ss_slow_time_boxcar(const double* in, int64_t rows, int64_t cols, int64_t window,
                        int64_t stride, int64_t n_out, double scale, double* out,
                        void* stream);

/* Straight-track time-domain backprojection through air (processing.py:
 * 270-348, focused_sar): field (n_traces, n_bins) complex128 interleaved,
 * twtt (n_bins) the bin delays (twtt[0] and twtt[1] - twtt[0] give t0 and
 * dt), s (n_traces) the non-decreasing along-track arc length, k the carrier
 * wavenumber 2 pi / lambda, window 0 rectangular / 1 hann; out (n_traces,
 * n_bins) complex128 interleaved (every element written).  SS_PREC_FP64
 * follows the reference expressions; SS_PREC_MIXED keeps the migrated delay
 * and phase cycles in float64 and interpolates and sums in float32. */
int ss_focused_sar(int precision, const double* field, int64_t n_traces, int64_t n_bins,
                   const double* twtt, const double* s, double c, double k, double aperture_m,
                   int window, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SOUNDERSIM_HIP_H */

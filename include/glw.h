/* glw -- MI355X (gfx950) Green--Laplace bound-wave reconstruction, C ABI.
 *
 * The reference has no FFI: its hot path is a NumPy/MATLAB call graph.  Each
 * entry point below replaces one reference function for one sea-state plan
 * (a fixed grid, parent support and GL rank set) and many time instants:
 *
 *   glw_plan_create / glw_plan_destroy
 *       replace the per-call setup of the reference graphs: supports, balance
 *       shifts, GL rules, multipliers (python/src/green_laplace/_order3.py:28-55,
 *       spatial.py:162-190, 241-275) and the MATLAB plan cache
 *       (src/internal/gl_no_stokes_eta33.m:535-621).  The host builds every
 *       FP64 table; the plan uploads them and owns the device workspace.
 *   glw_linear_execute
 *       eta11+ and psi11+ (spatial.py:74-79; src/gl_spectral_surface.m:26-31).
 *   glw_order2_execute
 *       eta22+ (spatial.py:156-231; src/internal/green_laplace_eta22.m) and
 *       psi22+ (spatial.py:234-414;
 *       src/internal/finite_depth_directional_dual_branch_green_laplace_psi22.m).
 *   glw_order3_execute
 *       eta33+ and psi33+ (_order3.py:18-148; src/internal/gl_no_stokes_eta33.m).
 *
 * Conventions: every array argument of an execute call is a device pointer.
 * Complex arrays are interleaved complex128.  Spectra are given on the parent
 * box (p1) of the plan, [p1.wy][p1.wx] row-major over signed mode numbers
 * (mx0.., my0..) of the ny x nx FFT grid.  Outputs are [nt][ny][nx] complex128
 * physical fields, row = y, already scaled to metres (eta) and m^2/s (psi),
 * identical in meaning to the reference's *_plus components.  Execute calls
 * never synchronise the host and run on `stream` (a hipStream_t, or NULL for
 * the default stream).  They allocate device memory in two cases only: the
 * split-row scratch, grown by the first eager call whose FFT schedule needs
 * it (refused under graph capture, and while a graph holds the plan,
 * glw_plan_hold), and the adjoint workspace on the first
 * glw_adjoint_execute.  Neither frees memory that an earlier call used
 * while a hold is active.  A plan must not be used
 * by two host threads or two streams at once.  Every function returns 0 on
 * success and a nonzero code on error; glw_last_error() returns the text
 * (thread-local).
 */
#ifndef GLW_H
#define GLW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLW_OK 0
#define GLW_EINVAL 1
#define GLW_EDEVICE 2
#define GLW_EUNSUPPORTED 3

#define GLW_PREC_FP64 0 /* FFTs and products in FP64 */
/* WIT only (glw_wit_desc.precision).  glw_plan_create refuses it with
 * GLW_EUNSUPPORTED: FP32 transforms of the GL graphs miss relative L2 1e-5
 * against FP64 for eta22, eta33, psi33 and eta44 (the node factors amplify
 * spectrum bins ~1e-7 below their maximum, which FP32 cannot resolve). */
#define GLW_PREC_FP32 1

/* Evaluation kinds (glw_plan_autotune, glw_batch_item). */
#define GLW_KIND_LINEAR 1 /* eta11, psi11 */
#define GLW_KIND_ORDER2 2 /* eta22, psi22 */
#define GLW_KIND_ORDER3 3 /* eta33, psi33 */
#define GLW_KIND_ORDER4 4 /* eta44 */

typedef struct glw_plan_s glw_plan;

/* Rectangle of signed mode numbers: mx0..mx0+wx-1, my0..my0+wy-1. */
typedef struct glw_box {
  int32_t mx0, my0, wx, wy;
} glw_box;

/* Host FP64 tables on a box, [wy][wx] row-major: qx=h*kx, qy=h*ky, nu =
 * sqrt(q tanh q), inu = 1/max(nu, DBL_MIN), q2 = q^2, a = q tanh q. */
typedef struct glw_box_tables {
  const double *qx, *qy, *nu, *inu, *q2, *a;
} glw_box_tables;

/* Order-2 stage tables (optional; order >= 2). */
typedef struct glw_order2_desc {
  int32_t eta_rank;             /* J of eta22 */
  const double* eta_damp;       /* [J][p1] exp(-tau_j (nu - c)) */
  const double* eta_cosh;       /* [J][p2] cosh(sqrt(A) tau_j) */
  const double* eta_sinhr;      /* [J][p2] sinh(sqrt(A) tau_j)/sqrt(A) */
  const double* eta_coef;       /* [J] w e^x / lambda / 4 * exp(-2 c tau) */
  int32_t channels;             /* psi22 radial channels (1 or 2) */
  const double* psi_damp;       /* [2 branches][2 nodes][channels][p1] channel mask * exp(-tau (nu - c_ch)) */
  const double* psi_factor;     /* [2][2][channels*channels][p2] output factor, zero off the channel-pair support */
  const double* psi_coef;       /* [2][2] w e^x / lambda / 8 */
  const double* psi_sign;       /* [2] sign of K/sqrt(A) per branch */
} glw_order2_desc;

/* Order-4 eta44 tables (optional; with_order4 != 0).  One GL rank J for every
 * nesting level, as paper/order4/cpp/gl_eta44_fftw.cpp uses. */
typedef struct glw_order4_desc {
  int32_t rank;                 /* J */
  glw_box p4;                   /* quartic box */
  int32_t g4x, g4y;             /* working grid of quartic products */
  glw_box_tables t4;            /* tables on p4 */
  const uint8_t* mask4;         /* [p4] quartic support */
  const double* damp_quartic;   /* [J][p1]     exp(-tau4 (nu - c4)) */
  const double* damp_third;     /* [J*J][p1]   damp_quartic[o4] exp(-tau3[o3] (nu - c3)) */
  const double* damp_pair3;     /* [J*J*J][p1] damp_third[o4,o3] exp(-tau2[i] (nu - c2)) */
  const double* damp_pair4;     /* [J*J][p1]   damp_quartic[o4] exp(-tau2[i] (nu - c2)) */
  const double *pair_cosh, *pair_sinhr, *pair_coef;          /* [J][p2], [J][p2], [J] */
  const double *third_cosh, *third_sinhr, *third_coef;       /* [J][p3], [J][p3], [J] */
  const double *quartic_cosh, *quartic_sinhr, *quartic_coef; /* [J][p4], [J][p4], [J] */
  /* 0: every forcing partition (1+1+1+1, 2+1+1, 2+2, 3+1); 1: the 3+1 partition
   * only.  With quartic_coef / third_coef zero outside one outer / third-order
   * node this gives the partial fields of the reference's split runs
   * (gl_eta44 ... OUTER4_INDEX [PART_INDEX], paper/order4/cpp/gl_eta44_fftw.cpp:639-716). */
  int32_t forcing_terms;
} glw_order4_desc;

typedef struct glw_plan_desc {
  int32_t nx, ny;               /* FFT grid */
  double depth, gravity;
  double floor_level;           /* 16 eps log2(nx ny): roundoff floor relative to the support max */
  int32_t precision;            /* GLW_PREC_* */
  int32_t max_instants;         /* workspace capacity (instants per internal batch) */
  int32_t order;                /* 1, 2 or 3 */
  glw_box p1, p2, p3;           /* parent, pair and triple boxes */
  int32_t g2x, g2y;             /* working grid of quadratic products */
  int32_t g3x, g3y;             /* working grid of cubic products */
  glw_box_tables t1, t2, t3;    /* tables on p1, p2, p3 */
  const double* omega;          /* [p1] angular frequency (rad/s) of the time phase */
  const uint8_t* mask2;         /* [p2] pair support */
  const uint8_t* mask3;         /* [p3] triple support */
  /* order 3 */
  int32_t inner_rank, outer_rank;
  const double* damp_outer;     /* [outer][p1] exp(-tau_o (nu - c3)) */
  const double* damp_inner;     /* [outer*inner][p1] exp(-tau_o (nu - c3)) exp(-tau_i (nu - c2)) */
  const double* inner_cosh;     /* [inner][p2] */
  const double* inner_sinhr;    /* [inner][p2] */
  const double* inner_coef;     /* [inner] w e^x / lambda2 * exp(-2 c2 tau) / 4 */
  const double* outer_cosh;     /* [outer][p3] */
  const double* outer_sinhr;    /* [outer][p3] */
  const double* outer_coef;     /* [outer] w e^x / lambda3 * exp(-3 c3 tau) */
  glw_order2_desc o2;
  int32_t with_order4;          /* build the eta44 graph (needs p2, p3, g2, g3 too) */
  glw_order4_desc o4;
} glw_plan_desc;

int glw_plan_create(const glw_plan_desc* desc, glw_plan** plan);
int glw_plan_destroy(glw_plan* plan);

/* eta11+, psi11+ at nt instants.  s0: [p1] complex128 analytic spectrum at
 * t = 0; times: [nt] seconds. */
int glw_linear_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt,
                       double* eta11, double* psi11, void* stream);

/* eta22+, psi22+ at nt instants (plan order >= 2). */
int glw_order2_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt,
                       double* eta22, double* psi22, void* stream);

/* eta33+, psi33+ at nt instants (plan order 3). */
int glw_order3_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt,
                       double* eta33, double* psi33, void* stream);

/* eta44+ at nt instants (plan with_order4): the pure-sum GL eta44 of
 * paper/order4/cpp/gl_eta44_fftw.cpp:639-752 and
 * src/internal/finite_depth_directional_green_laplace_eta44_three_kernels.m:57-219
 * (via gl_pure_sum_order4.m), scaled by depth. */
int glw_order4_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt, double* eta44,
                       void* stream);

/* WIT sector-44 quartet sum: replaces the per-quartet DirectionalKernel::evaluate
 * and the sector-44 loop of paper/order4/cpp/directional_eta4_openmp.cpp
 * (:157-453, :690-734, accumulate :595-602, output arrays :747-773).  Parent
 * subset states (singles, pairs, triples) are tabulated at creation. */
typedef struct glw_wit_desc {
  int32_t modes_count;          /* M */
  const int32_t* modes;         /* host [M][2] integer modes (mx, my) */
  int32_t nx, ny;               /* output grid */
  double gravity, depth, dkx, dky;
  double threshold;             /* near-resonance threshold on rho */
  int32_t precision;            /* GLW_PREC_FP64, or GLW_PREC_FP32: degree-4 evaluation in FP32 from FP32
                                   copies of the FP64 subset tables, FP64 weights and accumulation */
} glw_wit_desc;

typedef struct glw_wit_s glw_wit;

int glw_wit_create(const glw_wit_desc* desc, glw_wit** wit);
int glw_wit_destroy(glw_wit* wit);

/* amplitudes: device complex128 [cases][M].  out: device complex128
 * [5][cases][ny][nx] (x fastest), overwritten with the DIR44O01 arrays eta,
 * psi, eta coordinate, psi coordinate (zero in sector 44), retained.
 * counters: device double[8] = total, kept, near, strict-zero outputs,
 * strict-zero intermediates, max residual, min detuning, max strict-zero
 * source.  Accumulation uses FP64 atomics: bins are summed in a
 * schedule-dependent order (relative differences at the 1e-16 level). */
int glw_wit_sector44(glw_wit* wit, const double* amplitudes, int32_t cases, double* out, double* counters,
                     void* stream);
/* glw_wit_sector44 over the quartets whose first index i lies in
 * [first_begin, first_end) -- the iterations first_begin..first_end-1 of the
 * reference's outermost loop (directional_eta4_openmp.cpp:690-734 runs all
 * of them), and the unit a multi-GPU run would split the sum by.  `out` and
 * `counters` are overwritten with the range's sums; counters[0] (total) is
 * the range's quartet count.  flags: 0, or GLW_WIT_GLOBAL_ATOMICS to add
 * every quartet to `out` with global atomics instead of through the
 * workgroup's LDS window (the path windows above 128 KB take). */
#define GLW_WIT_GLOBAL_ATOMICS 1
int glw_wit_sector44_range(glw_wit* wit, int32_t first_begin, int32_t first_end, int32_t flags,
                           const double* amplitudes, int32_t cases, double* out, double* counters, void* stream);
/* Any sector of the reference runner (`--sector 40|42|44`,
 * directional_eta4_openmp.cpp:603-734): 44 is glw_wit_sector44; 40 (plus
 * pair x minus pair, analytic half plane) and 42 (plus triple x minus mode)
 * have signed parents, evaluated in FP64 with the retained-V3 push and the
 * y31 coordinate (:282-336).  Same arrays and counters as above (total = the
 * sector's candidate count).  The first call of sector 40 or 42 on a plan
 * allocates and fills the signed subset tables on `stream`: M^2 pair records
 * of 136 B and M^2 (M + 1) / 2 triple records of 168 B (675 MB at M = 200,
 * 43 GB at M = 800); later calls reuse them. */
int glw_wit_sector(glw_wit* wit, int32_t sector, const double* amplitudes, int32_t cases, double* out,
                   double* counters, void* stream);

/* Stage timing (instrumentation; no reference counterpart).  enable != 0
 * creates a pool of HIP events; every kernel launch of later execute calls is
 * bracketed by two events on the execute stream.  glw_plan_stage_times waits
 * for the last event, aggregates elapsed times per stage name and resets the
 * pool: names is filled with "name,name,..." (NUL-terminated), ms[i] and
 * launches[i] receive the total time and launch count of stage i, bytes[i]
 * (when bytes is not NULL) the stage's compulsory HBM bytes summed over its
 * launches -- every input element it consumes and output element it
 * produces on the reduced grids, once; tables shared by all tasks excluded
 * (0 for the adjoint's accumulations) -- and *count the number of stages
 * (at most max). */
int glw_plan_profile(glw_plan* plan, int32_t enable);
int glw_plan_stage_times(glw_plan* plan, int32_t max, char* names, size_t names_len, double* ms,
                         int32_t* launches, double* bytes, int32_t* count);

/* Adjoint (no reference counterpart: the reference has no gradients).  For
 * a real loss L of the elevation fields at one instant times[0], given
 * grad_field = dL/dRe(eta+) + i dL/dIm(eta+) (device complex128 [ny][nx],
 * the same for every included component), writes grad_s0 = dL/dRe S0 +
 * i dL/dIm S0 on the parent box (device complex128 [wy][wx] of glw_plan_desc
 * p1), where eta+ = sum of the components selected by `orders` (bit mask:
 * 1 eta11, 2 eta22, 4 eta33).  Supports, GL nodes, shifts, masks and
 * roundoff floors are those of the forward evaluation and are held fixed.
 * recompute != 0 first runs the forward stages for (s0, times[0]); with 0
 * the plan's workspace must still hold them (the last execute calls of
 * the selected orders were for this s0 and instant, nt = 1).  FP64 plans
 * only.  Allocates its workspace on the first call. */
int glw_adjoint_execute(glw_plan* plan, const double* s0, const double* times, int32_t orders,
                        const double* grad_field, double* grad_s0, int32_t recompute, void* stream);

/* FFT schedule autotuner (no reference counterpart: the reference runs
 * FFTW/MATLAB/NumPy FFTs whose planning is internal).  Runs `kind` of the
 * plan (reps timed repetitions per candidate, plus warm-ups) with every
 * candidate schedule -- radix sequence of the LDS FFTs (two-pass composite
 * radices or greedy radix 8/4/2/3) x tile (column tile width and threads per
 * workgroup; fused row kernel with the widest or a one-row tile, or the split
 * row path) -- times
 * every stage with HIP events and keeps, per stage, the fastest.  The
 * outputs a, b are overwritten.  Synchronizes the stream; not capturable. */
int glw_plan_autotune(glw_plan* plan, int32_t kind, const double* s0, const double* times, int32_t nt, double* a,
                      double* b, int32_t reps, void* stream);
/* The plan's per-stage schedule as "stage=plan/tile;..." (stages absent from
 * the text run plan 0 / tile 0), and the reverse.  Lets one tuning serve many
 * plans of the same grids. */
int glw_plan_get_schedule(glw_plan* plan, char* text, size_t len);
int glw_plan_set_schedule(glw_plan* plan, const char* text);
/* Hold (hold != 0) or release (hold == 0) a plan for captured graphs that
 * replay it (no reference counterpart: the reference has no graphs).  While
 * the hold count is positive, an execute that would need a larger split-row
 * scratch, and glw_plan_autotune, fail with GLW_EINVAL instead of moving
 * memory the graphs use, and glw_plan_destroy refuses the plan.
 * glw_batch_create holds each plan it captures until glw_batch_destroy. */
int glw_plan_hold(glw_plan* plan, int32_t hold);

/* Batched evaluation of many plans in one HIP graph (no reference
 * counterpart: the reference evaluates one sea state per call of
 * src/gl_spectral_surface.m / python/src/green_laplace/spatial.py).  Every
 * item is one execute call; glw_batch_create captures all of them into one
 * graph with `branches` parallel chains (items of one plan share a chain,
 * chains are balanced by estimated work), and glw_batch_launch replays the
 * whole batch with a single hipGraphLaunch.  Pointers are captured by value:
 * inputs are read and outputs written at every launch.  Each plan must have
 * run once eagerly with its current schedule (sizes its scratch) and must not
 * have stage profiling enabled. */
typedef struct glw_batch_item {
  glw_plan* plan;
  int32_t kind; /* GLW_KIND_* */
  int32_t nt;
  const double* s0;    /* device, as for the execute calls */
  const double* times; /* device double[nt] */
  double* a;           /* device complex128 [nt][ny][nx] */
  double* b;           /* device complex128 [nt][ny][nx] (unused for GLW_KIND_ORDER4) */
} glw_batch_item;

typedef struct glw_batch_s glw_batch;

int glw_batch_create(const glw_batch_item* items, int32_t n, int32_t branches, glw_batch** batch);
int glw_batch_launch(glw_batch* batch, void* stream);
int glw_batch_destroy(glw_batch* batch);

/* Device pointer and size of a named workspace buffer (for stage-level tests). */
int glw_plan_buffer(glw_plan* plan, const char* name, void** ptr, size_t* bytes);

/* Largest FFT length the kernels are instantiated for, and whether a length is one. */
int glw_fft_length_supported(int32_t n);

/* HBM calibration probe (measurement only; no counterpart in the reference):
 * times `reps` launches of a plain streaming kernel over `bytes` of device
 * memory on the default stream with HIP events and returns the mean launch
 * time in *ms.  kind 0: copy src -> dst; 1: fill dst (write only); 2: read src
 * (dst must still be a device buffer of at least 16 bytes: the read kernel
 * writes one value there only if a sum is infinite).  Synchronises. */
int glw_hbm_probe(int32_t kind, const double* src, double* dst, size_t bytes, int32_t reps, double* ms);

/* ---- Single-direction modal time path -------------------------------------
 *
 * glw_modal_execute replaces _modal_sum_components
 * (python/src/green_laplace/modal_time.py:187-341): eta22 and eta33 of one
 * level of the modal-points ladder, from a deposit of the first-order band
 * on the modal grid q_k = k dq.  The caller (host) selects the band, solves
 * the dispersion relation, brackets each q on the reference's P-point modal
 * axis (modal_time.py:344-368) and picks the reduced grids:
 *   q_len >= order * kspan + 1  (x points of the band-shifted modal axis),
 *   t_len >= order * span + 1   (t points of the band-shifted frequency axis),
 * lengths that glw_modal_length_supported accepts (bit 1: t_len, bit 2: q_len,
 * bit 4: a q_len transformed in one workgroup's LDS, up to 8192; the longer
 * ones, 12288 .. 65536, run as 2, 4 or 8 sub-transforms with a global
 * pre- and post-pass).
 * Input bins are base_bin ..
 * base_bin + span (work index w); deposit[w] = {Re A, Im A, w_lo, w_hi, k_lo}
 * (device, FP64, k_hi = k_lo + 1, k0 = min k_lo, kspan = max k_lo + 1 - k0).
 * Outputs (device): eta11 (double[count], may be NULL) = Re fft of the
 * positive-bin amplitudes, eta22 and eta33 (complex128[count], the analytic
 * components; eta33 only for order 3, may be NULL otherwise), audit
 * (double[4], may be NULL) = {eta22 invalid, eta22 total, eta33 invalid,
 * eta33 total} source energy.  The workspace (device, `workspace_bytes` >=
 * glw_modal_workspace_size) is used as scratch; nothing is allocated. */
typedef struct glw_modal_desc {
  int32_t order;           /* 2 or 3 */
  int32_t count;           /* samples N of the record */
  int32_t base_bin, span;  /* input bins base_bin .. base_bin + span */
  int32_t q_len, t_len;    /* reduced grid lengths */
  int32_t k0, kspan;       /* deposit band k0 .. k0 + kspan */
  int32_t rank;            /* Gauss--Laguerre rank (<= 16) */
  double dq;               /* modal spacing (dimensionless wavenumber) */
  double fstep;            /* 2 pi / (N dt) sqrt(h / g) */
  double depth;
  double lambda2;          /* eta22 scale sqrt(4 qp tanh qp - 2 qp tanh 2qp) */
  double lambda_inner;     /* pair scale 2 nu_p - sqrt(2 qp tanh 2qp) */
  double lambda3;          /* eta33 scale 3 nu_p - sqrt(3 qp tanh 3qp) */
  double nodes[16], weights[16];
} glw_modal_desc;

int glw_modal_length_supported(int32_t n);
int glw_modal_workspace_size(const glw_modal_desc* desc, size_t* bytes);
int glw_modal_execute(const glw_modal_desc* desc, const double* deposit, void* workspace, size_t workspace_bytes,
                      double* eta11, double* eta22, double* eta33, double* audit, void* stream);
/* glw_modal_execute bracketed by HIP events; waits for them and returns the
 * durations (ms[6]: twiddle tables, order-2 time pass, eta22 modal pass,
 * order-3 time pass, eta33 modal pass, output synthesis).  Measurement only:
 * no counterpart in the reference. */
int glw_modal_stage_times(const glw_modal_desc* desc, const double* deposit, void* workspace, size_t workspace_bytes,
                          double* eta11, double* eta22, double* eta33, double* audit, void* stream, float* ms);
/* glw_modal_execute without a workspace argument, allocating nothing (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops).  The
 * workspace is a static 128 MiB device array the library holds on every GPU
 * it runs on until the process ends; levels whose glw_modal_workspace_size
 * exceeds it fail with GLW_EINVAL.  Calls on one GPU share the array: calls
 * that overlap on one GPU race. */
int glw_modal_execute_noalloc(const glw_modal_desc* desc, const double* deposit, double* eta11, double* eta22,
                              double* eta33, double* audit, void* stream);

/* glw_eta20_execute replaces unidirectional_eta20
 * (python/src/green_laplace/_eta20.py:13-132): the difference-frequency
 * elevation on the reference's nx x ny modal grid, evaluated on its qy = 0
 * plane (the unidirectional input deposits nothing else).  The deposit
 * brackets each q on the eta20 qx axis (nx points, _eta20.py:44,156-172);
 * q_len >= 2 kspan + 1 (an LDS length: bit 4 of glw_modal_length_supported),
 * t_len >= 2 span + 1.  Output eta20 (device,
 * double[count]) after the Hermitian projection (_eta20.py:175-185).
 * delta_q is the GL scale (_eta20.py:90-93); the caller returns zeros for
 * the degenerate case. */
typedef struct glw_eta20_desc {
  int32_t count, base_bin, span, q_len, t_len, k0, kspan, rank;
  double dq, fstep, depth, delta_q;
  double nodes[16], weights[16];
} glw_eta20_desc;

int glw_eta20_workspace_size(const glw_eta20_desc* desc, size_t* bytes);
int glw_eta20_execute(const glw_eta20_desc* desc, const double* deposit, void* workspace, size_t workspace_bytes,
                      double* eta20, void* stream);

const char* glw_last_error(void);
const char* glw_version(void);

/* ABI self-check (no reference counterpart: the reference has no C ABI).
 * sizeof the public descriptor struct `which` as this library was compiled,
 * so a binding can compare its own layout before passing one (0: unknown). */
enum {
  GLW_ABI_BOX = 1, GLW_ABI_BOX_TABLES = 2, GLW_ABI_ORDER2_DESC = 3, GLW_ABI_ORDER4_DESC = 4,
  GLW_ABI_PLAN_DESC = 5, GLW_ABI_WIT_DESC = 6, GLW_ABI_BATCH_ITEM = 7, GLW_ABI_MODAL_DESC = 8,
  GLW_ABI_ETA20_DESC = 9
};
size_t glw_abi_sizeof(int32_t which);

#ifdef __cplusplus
}
#endif

#endif /* GLW_H */

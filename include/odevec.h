/*
 * libodevec -- C ABI of the MI355X-native OdeVec BDF integrator.
 *
 * The reference's host interface is the Fortran module `odevec_main`
 * (src/odevec.F90): a `type(odevec)` handle with caller-set tolerance fields,
 * `InitOdevec(this)`, `SolveODE(this,time,dt,t_stop,y,GetRHS,GetJac,GetLU,Mask)`
 * and `CloseOdevec(this)`.  Each entry point below names the reference
 * interface it replaces.  All sizes are plain integers, all arrays plain
 * pointers; no other library's types appear.
 *
 * State layout: y is the reference's Fortran y(nvector,neq) -- neq planes of
 * ncells doubles, cells fastest: y[j*ncells + i].  Parameters likewise:
 * params[p*ncells + i].
 *
 * Error handling: every call returns ODEVEC_OK (0) or a negative code;
 * odevec_last_error() gives the message of the calling thread's last error.
 * Per-cell solver failures (the reference prints and STOPs,
 * src/odevec.F90:380-389) are reported as per-cell status codes and make the
 * call return ODEVEC_ESOLVE; with opts.abort_on_error the library prints and
 * aborts instead, like the reference.
 *
 * Threading: a handle owns its device buffers and HIP stream; use one host
 * thread per handle.  There is no global mutable state besides the
 * thread-local error message and the read-only network registry.
 */
#ifndef ODEVEC_H
#define ODEVEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ODEVEC_ABI_VERSION 2

/* return codes */
#define ODEVEC_OK 0
#define ODEVEC_EINVAL (-1)    /* bad argument */
#define ODEVEC_ENOMEM (-2)    /* device or host allocation failed */
#define ODEVEC_EDEVICE (-3)   /* HIP runtime error */
#define ODEVEC_ENOTFOUND (-4) /* unknown network id */
#define ODEVEC_ESOLVE (-5)    /* at least one cell did not finish with status OK/SKIPPED */

/* per-cell status codes */
#define ODEVEC_STATUS_OK 0
#define ODEVEC_STATUS_DT_MIN 1     /* step below MinimalTimestep (src/odevec.F90:272-275, 387-390) */
#define ODEVEC_STATUS_CONV_FAIL 2  /* > 10 consecutive convergence failures (src/odevec.F90:380-383) */
#define ODEVEC_STATUS_ORDER 3      /* order/step mismatch (src/odevec.F90:584-587, 595-598) */
#define ODEVEC_STATUS_MAX_STEPS 4  /* attempt budget exhausted (no reference counterpart) */
#define ODEVEC_STATUS_SKIPPED 5    /* masked out or t_stop <= t0: y unchanged */
#define ODEVEC_STATUS_PAUSED 8     /* batch segment with a quantum: not finished in this launch (y unchanged) */
#define ODEVEC_STATUS_FRESH 9      /* set by the caller before a quantum segment's first launch */

/* precision forms */
#define ODEVEC_FP64 0   /* everything FP64 */
#define ODEVEC_MIXED 1  /* P, LU factors and triangular solves FP32; residual, history, norms, control FP64 */

/* execution modes */
#define ODEVEC_PERCELL 0     /* one lane per cell, per-cell step control (reference at nvector=1) */
#define ODEVEC_LOCKSTEP 1    /* groups of `group` cells share dt/order/verdicts (reference at nvector=group) */
#define ODEVEC_PERSISTENT 2  /* per-cell control, persistent grid with lane refill (work compaction) */
#define ODEVEC_COOP 3        /* per-cell control, one cell per group of lanes with one species per lane
                                (CSC networks; the only mode for networks of more than 32 species) */

typedef struct odevec_s odevec_t;

typedef struct {
  int32_t precision;      /* ODEVEC_FP64 | ODEVEC_MIXED */
  int32_t mode;           /* ODEVEC_PERCELL | ODEVEC_LOCKSTEP | ODEVEC_PERSISTENT */
  int32_t group;          /* lockstep group size, 1..256; 0 = min(nvector, 256) */
  int32_t block;          /* threads per workgroup (percell/persistent), 0 = 64 */
  int32_t grid;           /* persistent: workgroups, 0 = fill the device */
  int32_t max_attempts;   /* per-cell step-attempt budget, 0 = 1000000 */
  int32_t device_ptrs;    /* 1: array arguments are device pointers (no copies) */
  int32_t abort_on_error; /* 1: print and abort on a failed cell, like the reference's STOP */
  int32_t schedule;       /* persistent: 0 = index order, 1 = costliest cells first (ranked by
                             log2(span / initial step) before the launch); results are the same */
} odevec_opts_t;

typedef struct {
  int64_t attempts, steps, newton, jac, errfail, convfail, stepsel, qsum; /* summed over cells */
  int64_t cells_ok, cells_failed, cells_skipped;
  double flops64, flops32; /* algorithmic flops of the solve (op-count model) */
  double kernel_ms;        /* device time of the solve kernel(s) */
} odevec_stats_t;

typedef struct {
  int32_t neq, nnz, nrate, nparam, rate_kind, tgas_index;
  int32_t flops_rhs, flops_jac, flops_lu, flops_solve, flops_lu_r, flops_solve_r;
} odevec_network_info_t;

typedef struct {
  int32_t vgprs, spill_bytes, blocks_per_cu, block;
} odevec_kernel_info_t;

/* Library / registry ------------------------------------------------------ */
int32_t odevec_abi_version(void);
const char* odevec_last_error(void);
int32_t odevec_network_count(void);
const char* odevec_network_name(int32_t index);           /* e.g. "PRIMORDIAL/CMK/CSC" */
int32_t odevec_network_info(const char* network, odevec_network_info_t* info);
int32_t odevec_kernel_info(const char* network, int32_t mode, int32_t precision, int32_t block,
                           odevec_kernel_info_t* info);
void odevec_default_opts(odevec_opts_t* opts);

/* Replaces InitOdevec(this) (src/odevec.F90:94-154): a handle for `nvector`
 * cells of the generated network `network` (pre_odevec.py --example /
 * --ordering / --packaging select it). opts may be NULL (defaults). */
int32_t odevec_create(const char* network, int64_t nvector, const odevec_opts_t* opts, odevec_t** out);

/* Replace the caller-set fields RelativeTolerance, AbsoluteTolerance,
 * MinimalTimestep and CheckNegatives of type(odevec) (src/odevec.F90:48-52).
 * Defaults: rtol 1e-6, atol 1e-10, dt_min 1e-10 (single-precision literal as
 * pre_odevec.py:258-259 emits it), check_negatives 0 (src/odevec.F90:146). */
int32_t odevec_set_tolerances(odevec_t* h, double rtol, double atol);
int32_t odevec_set_dt_min(odevec_t* h, double dt_min);
int32_t odevec_set_check_negatives(odevec_t* h, int32_t flag);

/* FP64 fallback of the mixed form (no reference counterpart). With
 * convfail_limit > 0, a cell (per-cell or persistent mode) that reaches
 * convfail_limit convergence failures stops after that
 * failure's normal handling (history reset, step size x 0.25, a Jacobian
 * refresh due), leaves its integrator state in a hand-off record and keeps
 * its input state; FP64 lanes continue it from there. In persistent mode they
 * run beside the launch, on a second stream of the handle (forked from and
 * joined back into `stream`, in 1 in 32 of the machine's wavefront slots),
 * taking each cell as it is handed back; in per-cell mode, and for whatever
 * they leave, the FP64 persistent kernel continues the cells after the launch
 * on the same stream. Counters are those of both precisions together. A few
 * cells fall into long runs of convergence failures with FP32 factors (tens
 * of thousands of attempts where FP64 needs about a thousand) and would set
 * the whole launch's length; the fallback bounds them. On an FP64 handle the
 * cells only move between lanes: results are those without the fallback, bit
 * for bit. 0 (default): off. */
int32_t odevec_set_fp64_fallback(odevec_t* h, int32_t convfail_limit);

/* Invariant test of the mixed form's corrector (no reference counterpart):
 * in the mixed form a corrector iteration that passes the convergence test
 * but leaves one of the network's linear invariants c (element and charge
 * conservation, sum_j c_j f_j = 0) with a residual |c^T (y_NS(:,1) + en)|
 * above `tolerances` x sqrt(sum_j c_j^2 (rtol |y_j| + atol)^2) counts as a
 * convergence failure (a quarter step and fresh factors, the reference's
 * path at src/odevec.F90:375-393). Exact factors of P = I - dt*beta*J keep
 * the invariants at every corrector step (c^T P = c^T); FP32 factors of a
 * very stiff P do not, and without the test the cell's invariants drift.
 * With the FP64 fallback on (odevec_set_fp64_fallback), a cell whose test
 * fails `fallback_fails` times is handed to the FP64 pass like one that
 * reaches the convergence-failure limit (the two counts are separate);
 * without a fallback the test is no longer applied to that cell after
 * `fallback_fails` failures, which bounds the work of a cell whose FP32
 * factors keep failing it (0: neither). Defaults ODEVEC_MIXED_INVARIANT_TOL and
 * ODEVEC_FALLBACK_INVARIANT_FAILS; tolerances 0 turns the test off. The FP64
 * form and networks without linear invariants (OREGO) are unaffected. */
#define ODEVEC_MIXED_INVARIANT_TOL 1.0
#define ODEVEC_FALLBACK_INVARIANT_FAILS 32
int32_t odevec_set_invariant_test(odevec_t* h, double tolerances, int32_t fallback_fails);

/* Convergence test of the mixed form's corrector (no reference counterpart;
 * the reference's CheckConvergence is src/odevec.F90:622-665).
 * ODEVEC_MIXED_TEST_MAX (the default): every weighted component of the last
 * correction within the divisor, with no rate discount -- FP32 factors of a
 * very stiff P contract trace species and invariants more slowly than the
 * RMS norm shows (PRIMORDIAL). ODEVEC_MIXED_TEST_RMS: the reference's test,
 * the RMS norm times min(1, 1.5 rate), as the FP64 form uses it -- for
 * networks whose FP32 factors contract every component (SYNTH64: the same
 * corrector iterations as FP64, 35% fewer than under the max norm, at the
 * same distance to an rtol-1e-12 solution). The FP64 form is unaffected. */
#define ODEVEC_MIXED_TEST_MAX 0
#define ODEVEC_MIXED_TEST_RMS 1
int32_t odevec_set_mixed_convergence_test(odevec_t* h, int32_t test);

/* Replaces SolveODE(this,time,dt,t_stop,y,GetRHS,GetJac,GetLU,Mask)
 * (src/odevec.F90:158-214) for the handle's nvector cells: integrate y from
 * *time to t_stop; *dt is the initial-step hint on entry and the last step
 * size (cell 0's in per-cell modes) on return; *time returns as t_stop.
 * mask (NULL = all cells) and status (NULL = not wanted) have nvector
 * entries. The network's RHS/Jacobian/LU are fixed at generation time, so
 * the reference's procedure arguments have no counterpart here. */
int32_t odevec_solve(odevec_t* h, double* time, double* dt, double t_stop, double* y,
                     const int8_t* mask, int32_t* status);

/* Per-cell form (no reference counterpart; the reference shares one t_stop
 * per call): cell i integrates from t0[i] to t_stop[i] with dt hint dt0[i].
 * t0 NULL -> 0, dt0 NULL -> 1e-10, params NULL -> network defaults, status /
 * dt_out / stats NULL -> not wanted. ncells may exceed the handle's nvector
 * (buffers grow). */
int32_t odevec_solve_cells(odevec_t* h, int64_t ncells, const double* t0, const double* t_stop,
                           const double* dt0, double* y, const double* params, int32_t* status,
                           double* dt_out, odevec_stats_t* stats);

/* Asynchronous per-cell form on a caller stream (a hipStream_t passed as
 * void*; NULL = the handle's stream). All pointers are device pointers;
 * strides of 0 broadcast one value to every cell (params: stride between
 * cells, planes are ncells apart when the cell stride is 1, else 1 apart).
 * counters (optional) receives 8 int32 planes of ncells: attempts, steps,
 * newton, jac, errfail, convfail, stepsel, qsum. flops (optional) receives
 * 2 doubles (FP64, FP32), accumulated. The call returns after enqueueing. */
int32_t odevec_solve_cells_async(odevec_t* h, int64_t ncells, const double* d_t0, int64_t t0_stride,
                                 const double* d_t_stop, int64_t t_stop_stride, const double* d_dt0,
                                 int64_t dt0_stride, double* d_y, const double* d_params,
                                 int64_t params_cell_stride, int32_t* d_status, double* d_dt_out,
                                 int32_t* d_counters, double* d_flops, void* stream);

/* Per-cell solve that allocates nothing. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * solve of odevec_solve_cells_async in persistent mode (cooperative mode for
 * networks of more than 32 species), without a handle: every device buffer
 * beyond the caller's arrays -- the queue counter, the per-wavefront
 * workspace and the default values -- is a fixed file-scope device array of
 * ODEVEC_NOALLOC_SCRATCH_BYTES, one per GPU, held until the process ends.
 * All pointers are device pointers with one value per cell (d_t0 NULL -> 0,
 * d_dt0 NULL -> 1e-10, d_params NULL -> the network's defaults, d_status /
 * d_dt_out NULL -> not wanted); the call enqueues on `stream` (a hipStream_t)
 * and returns. grid: workgroups, 0 = fill the device or the largest grid
 * whose workspace fits the scratch; a grid whose workspace does not fit, or
 * ncells >= 2^31, is rejected with ODEVEC_EINVAL. No FP64 fallback, no cost
 * order, check_negatives off, the default invariant test. Every call on one
 * GPU shares that GPU's scratch: calls that overlap on one GPU (on different
 * streams, or from different threads) race; calls in one stream are
 * ordered. Results equal odevec_solve_cells_async's in persistent mode bit
 * for bit. */
#define ODEVEC_NOALLOC_SCRATCH_BYTES (24u << 20)
int32_t odevec_solve_cells_noalloc(const char* network, int32_t precision, int64_t ncells, const double* d_t0,
                                   const double* d_t_stop, const double* d_dt0, double* d_y, const double* d_params,
                                   int32_t* d_status, double* d_dt_out, double rtol, double atol, double dt_min,
                                   int32_t grid, void* stream);

/* The generated plugin's network functions, evaluated on the GPU for
 * ncells cells of the handle's network, in the reference's layouts (arrays
 * are host arrays, or device arrays with opts.device_ptrs):
 *   GetRHS(this, y, rhs) (src/odevec_commons.F90:78-95):
 *     rhs[j*ncells + i] = f_j(y_i) for all neq species (frozen ones 0);
 *   GetJac(this, beta, y, dt, jac) (src/odevec_commons.F90:32-51,
 *     pre_odevec.py:144-180): P = I - beta*dt*J (permuted) in the
 *     packaging's storage, jac[e*ncells + i] for e < odevec_jac_entries:
 *     DENSE e = c*neq + r holds P_order(r, c) (Fortran jac(:, r, c)); CSC e
 *     is the e-th entry of the LU pattern (jac%sdata(:, e), fill entries 0).
 * Rates come from each cell's y and params (NULL = the network defaults).
 * The reference's GetLU (LUmethod 2) has no counterpart. */
int32_t odevec_get_rhs(odevec_t* h, int64_t ncells, const double* y, const double* params, double* rhs);
int32_t odevec_get_jac(odevec_t* h, int64_t ncells, double beta, double dt, const double* y, const double* params,
                       double* jac);
int32_t odevec_jac_entries(const char* network, int64_t* entries);

/* Replaces CloseOdevec(this) (src/odevec.F90:1058-1074). */
int32_t odevec_destroy(odevec_t* h);

/* Warm restart --------------------------------------------------------------
 * No reference counterpart: the reference's SolveODE cold-starts every call
 * (src/odevec.F90:170-182: order 1, fresh Jacobian, initial step size). With
 * a history attached, every per-cell solve on the handle (sync or async,
 * per-cell, persistent or cooperative mode) saves each successfully solved cell's
 * integrator state (Nordsieck history, order, step size, LU factors, error
 * weights, rates, Jacobian-refresh counters) into the history, and a later
 * call continues a cell from it when the call starts at the saved t_stop on
 * the state it returned (t0 and every species and parameter bitwise equal;
 * any other cell starts cold). A solve to t1 continued to t2 then takes the
 * same steps as one solve to t2 and returns the same result bit for bit,
 * with the FP64 fallback too: the convergence- and invariant-failure counts
 * the fallback decides on carry over (their counters then count from the
 * cell's first call), and a cell that FP64 lanes continued is continued by
 * FP64 lanes again, from its history.
 *
 * The cooperative mode keeps histories (in a layout of its own) for networks
 * too large for one lane, which only it runs (SYNTH64); for the others it
 * refuses one.
 *
 * odevec_history_planes: doubles per cell of a history for `network`.
 * odevec_set_history: attach d_hist, a device array of planes x ncells
 * doubles (cell i's record at d_hist[i*planes .. (i+1)*planes)) that the caller
 * allocates zero-filled and keeps; every solve must then cover exactly
 * ncells cells. NULL detaches. Calls sharing one history must not overlap. */
int32_t odevec_history_planes(const char* network, int64_t* planes);
int32_t odevec_set_history(odevec_t* h, double* d_hist, int64_t ncells);

/* Timeline (diagnostics; no reference counterpart): with d_times attached
 * (device, 2 x ncells int64), every solve on the handle writes the device
 * wall-clock tick at which each cell was loaded (d_times[i]) and finished
 * (d_times[ncells + i]); ticks_per_second (optional) receives the clock's
 * rate. It shows a launch's tail: how long the last cells run after the rest
 * are done. Per-cell, persistent and cooperative modes; every solve must
 * cover exactly ncells cells; NULL detaches. */
int32_t odevec_set_timeline(odevec_t* h, int64_t* d_times, int64_t ncells, int64_t* ticks_per_second);

/* Batches ------------------------------------------------------------------
 * No reference counterpart: the reference solves one network per build and
 * call. A bundle is one persistent kernel compiled for several (network,
 * precision) entries; a batch launch runs a table of segments -- blocks of
 * cells, each of one entry, with its own arrays, tolerances and queue -- in
 * ONE launch (work compaction across the whole batch). Every cell keeps its
 * own step control, so a cell's result does not depend on the other
 * segments (bitwise, for a given build). Segments are processed in table
 * order; put the most expensive per cell first. */
typedef struct odevec_batch_s odevec_batch_t;

typedef struct {
  const char* network;      /* variant id, one of the bundle's entries, e.g. "PRIMORDIAL/CMK/CSC" */
  int32_t precision;        /* ODEVEC_FP64 | ODEVEC_MIXED (must be compiled in the bundle) */
  int32_t check_negatives;
  int64_t ncells;
  double* y;                /* device: [neq][ncells], in/out */
  const double* t0;         /* device: [ncells], NULL = 0 */
  const double* t_stop;     /* device: [ncells] */
  const double* dt0;        /* device: [ncells], NULL = 1e-10 */
  const double* params;     /* device: [nparam][ncells], NULL = the network's defaults */
  int32_t* status;          /* device: [ncells] or NULL */
  double* dt_out;           /* device: [ncells] or NULL */
  int32_t* counters;        /* device: [8][ncells] or NULL */
  double* flops;            /* device: [2] accumulated (FP64, FP32) or NULL */
  double rtol, atol, dt_min;
  int32_t max_attempts;     /* 0 = 1000000 */
  int32_t quantum;          /* > 0 (needs hist and status, no flops): a cell runs at most this many
                               step attempts per launch, then pauses with status ODEVEC_STATUS_PAUSED
                               and its state in hist; the next launch of the same segment (same arrays,
                               start and end times) continues it. Set status to ODEVEC_STATUS_FRESH
                               before the first launch; cells with a final status are left as they
                               are. The result equals the solve in one launch bit for bit. hist then
                               holds pause records only (needs no zero fill; no warm restart). 0: off */
  double* hist;             /* device: warm-restart history of these cells (odevec_history_planes x ncells
                               doubles, zero-filled before first use) or NULL = cold start; see
                               odevec_set_history for the rules */
} odevec_segment_t;

int32_t odevec_bundle_count(void);
const char* odevec_bundle_name(int32_t index);   /* e.g. "MIX3", "MIX3+REPRO" */
/* Entry `index` of a bundle: its network id and precision; returns the
 * number of entries (or a negative code). */
int32_t odevec_bundle_entry(const char* bundle, int32_t index, const char** network, int32_t* precision);
/* block: threads per workgroup (0 = 256); grid: workgroups (0 = fill the
 * device); use_graph: capture the launch (queue reset + batch kernel) in a
 * hipGraph once per launch shape and stream, and replay it. */
int32_t odevec_batch_create(const char* bundle, int32_t block, int32_t grid, int32_t use_graph, odevec_batch_t** out);
/* Enqueue one launch over nseg (<= 1024) segments on `stream` (a
 * hipStream_t). Use one stream per batch handle: consecutive calls reuse its
 * segment table, queue counters and workspace in stream order. */
int32_t odevec_batch_solve_async(odevec_batch_t* b, const odevec_segment_t* segs, int32_t nseg, void* stream);

/* FP64 fallback for the batch's mixed segments (no reference counterpart;
 * see odevec_set_fp64_fallback): with convfail_limit > 0, a cell of a mixed
 * segment (with or without history) that reaches convfail_limit convergence failures
 * is handed off as for a handle; one FP64 kernel per such segment runs beside
 * the batch launch on the batch's second stream (outside the captured graph)
 * and continues the segment's cells as they are handed back, and the
 * network's FP64 persistent kernel takes what they left after the launch.
 * 0 (default): off. */
int32_t odevec_batch_set_fp64_fallback(odevec_batch_t* b, int32_t convfail_limit);
/* The invariant test of the batch's mixed segments (see odevec_set_invariant_test). */
int32_t odevec_batch_set_invariant_test(odevec_batch_t* b, double tolerances, int32_t fallback_fails);
int32_t odevec_batch_kernel_info(const char* bundle, int32_t block, odevec_kernel_info_t* info);
int32_t odevec_batch_destroy(odevec_batch_t* b);

#ifdef __cplusplus
}
#endif
#endif /* ODEVEC_H */

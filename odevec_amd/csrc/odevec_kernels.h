// HIP kernels around the per-lane integrator.
//
//  * odv_k_percell     -- one lane per cell, one cell per lane (the plain form);
//  * odv_k_persistent  -- a grid sized to the machine; a lane that finishes its
//                         cell takes the next one from a queue (one atomic per
//                         wavefront per refill round), so lanes whose cells need
//                         few steps do not idle while the wavefront's slowest
//                         cell finishes;
//  * odv_k_lockstep    -- one workgroup per group of `group` cells sharing one
//                         step size, order and verdicts: the reference's
//                         semantics at nvector = group (max-norms over the
//                         group, src/odevec.F90:669-702).
//
// Data layout in HBM (struct of arrays, cells fastest, as the reference's
// Fortran y(nvector,neq)): y[j*ncells + i], params[p*ncells + i].
// For networks with neq >= 8 (the per-cell kernels) part of a lane's state
// lives outside its registers, as columns of its wavefront (element r of
// lane t at [r*64 + t] in the wavefront's LDS slice: consecutive lanes hit
// consecutive banks), placed by LdsPolicy per precision. Each wavefront owns
// its slice, so a workgroup of any size (64..256 lanes) is a set of
// independent wavefronts and a CU refills one wavefront's slot as soon as
// that wavefront ends. The register file (512 VGPR+AGPR per lane at one
// wavefront per SIMD) and LDS (640 B per lane at 256 lanes per CU) hold
// ~2.6 KiB per lane together:
//  * mixed form (LU factors FP32, 4 B each): Nordsieck columns 0, 1, 5 and
//    the error weights stay in registers; LDS holds the rate coefficients
//    (read by every RHS and Jacobian) and Nordsieck columns 2..4; the
//    previous correction (read once per step-size selection) goes to the
//    device workspace;
//  * FP64 form (LU factors 8 B each): LDS holds the Nordsieck columns
//    2..MAXORD, the error weights and the previous correction; the rate
//    coefficients go to a device workspace with the same layout per
//    wavefront slot (ws[slot*WS_PER_SLOT + r*64 + t]; 512 B contiguous per
//    rate and slot, so a slot's rates stay in the L2 of the XCD running it).
#pragma once
#include "odevec_integrator.h"

namespace odv {
inline namespace ODV_VARIANT_NS {

struct KArgs {
  double* y;              // [NEQ][ncells], in/out
  long long ncells;
  const double* prm;      // params: prm[p*ps_plane + i*ps_cell]
  long long ps_plane, ps_cell;
  const double* t0;       // t0[i*t0_s]
  const double* tstop;    // tstop[i*ts_s]
  const double* dt0;      // dt hint dt0[i*dt_s]
  long long t0_s, ts_s, dt_s;
  const signed char* mask;  // optional
  int* status;            // optional [ncells]
  double* dt_out;         // optional [ncells]: last step size
  int* counters;          // optional [8][ncells]
  double* flops;          // optional [2]: FP64, FP32 algorithmic flops (atomic)
  unsigned long long* queue;  // persistent: next cell (set to the number of launched lanes)
  double* ws;             // rate workspace, LdsPolicy::WS_PER_SLOT doubles per wavefront slot
  double* hist;           // optional warm-restart history: Hist<Net>::PLANES doubles per cell, cell-major
  const int* order;       // persistent: queue position -> cell (costliest first), or null = index order
  long long* times;       // optional [2][ncells]: device wall-clock ticks when each cell started / ended
  int* promo;             // ctl.promote_cf > 0: cells stopped with ST_PROMOTE are appended here ...
  unsigned* promo_n;      // ... at promo[atomicAdd(promo_n, 1)]; their y, status, counters are not written
  const unsigned* qlen;   // persistent: number of queue positions (device count), or null = ncells
  // FP64 fallback hand-off records, Handoff<Net>::REC doubles per cell:
  // ho_mode 1 -- a cell stopped with ST_PROMOTE leaves its state in its
  // record; ho_mode 2 -- the cells of the queue continue from theirs (FP64)
  double* ho;
  int ho_mode;
  // persistent mode: a wavefront that leaves the loop adds one here once its
  // list entries are stored (the FP64 lanes of a concurrent kernel wait for
  // all of them before declaring the list complete)
  unsigned* fb_done;
  Controls ctl;
  int group;              // lockstep group size
  // persistent mode without an order: the arrays are already in queue order
  // (k_gather), and lanes start spread over the wavefronts as with an order
  int spread;
  // > 0 (persistent and batch launches with a history): a cell runs at most
  // this many attempts per launch, then pauses (ST_PAUSED, its state in its
  // history) for the next launch of the same solve; status[] carries each
  // cell's phase between the launches (ST_FRESH before the first). The
  // history then holds pause records only (no warm restart; -1: the FP64
  // fallback's passes over such a solve, which run cells to the end)
  int quantum;
};

constexpr int WAVE = 64;         // lanes per wavefront: the column stride of a lane's external state
constexpr int BLOCK = 64;        // default lanes per workgroup of the per-cell kernels
constexpr int MAX_BLOCK = 256;   // largest workgroup of the per-cell kernels

constexpr int LDS_BYTES = 160 * 1024;   // LDS per CU on gfx950
// persistent kernel: lanes waiting for a new cell before a wavefront refills
// (1M-cell PRIMORDIAL, mixed with the FP64 fallback: 1 -> 221 ms, 4 -> 212 ms,
// 8 -> 213 ms per solve; results do not depend on it)
constexpr int REFILL_MIN = 4;
constexpr int LDS_WAVES = 4;            // resident wavefronts per CU the LDS placement is sized for

template <class Net, class M>
struct LdsPolicy {
  static constexpr int N = Net::NACT;
  static constexpr bool FP64 = sizeof(M) == sizeof(double);
  // LDS doubles per lane -- mixed: rates, Nordsieck columns 2..4; FP64:
  // Nordsieck columns 2..MAXORD, weights, previous correction
  static constexpr int COLS = FP64 ? (MAXORD - 1) * N + 2 * N : Net::NRATE + 3 * N;
  static constexpr bool USE = Net::NEQ >= 8 && COLS * 8 * WAVE * LDS_WAVES <= LDS_BYTES;
  static constexpr int KS = USE ? WAVE : 0;
  static constexpr int ZX = USE ? (FP64 ? MAXORD - 1 : 3) : 0;   // Nordsieck columns from 2 on in LDS
  static constexpr bool KLDS = USE && !FP64;       // rates in LDS (else in the workspace)
  static constexpr bool WX = USE && FP64;          // error weights in LDS
  static constexpr bool ELDS = USE && FP64;        // previous correction in LDS (else in the workspace)
  static constexpr int KWORDS = KLDS ? Net::NRATE * WAVE : 0;
  static constexpr int ZWORDS = ZX > 0 ? ZX * N * WAVE : 0;
  static constexpr int WWORDS = WX ? N * WAVE : 0;
  static constexpr int EWORDS = ELDS ? N * WAVE : 0;
  static constexpr long long WS_RATES = (USE && !KLDS) ? (long long)Net::NRATE * WAVE : 0;
  static constexpr long long WS_PER_SLOT = WS_RATES + ((USE && !ELDS) ? (long long)N * WAVE : 0);   // per wavefront
  static constexpr int WORDS = KWORDS + ZWORDS + WWORDS + EWORDS;   // one wavefront's LDS slice
};

extern __shared__ double odv_lane_lds[];   // the workgroup's wavefront slices (dynamic size)

// LDS bytes of a workgroup of `block` lanes whose wavefronts take `words` doubles each
inline constexpr int lane_lds_bytes(int words, int block) { return words * 8 * (block / WAVE); }

// Point a lane's external columns at its wavefront's LDS slice `lds` and at
// its wavefront's slot of the rate workspace (slot = global wavefront id).
template <class P, class L>
__device__ inline void odv_bind_columns(L& s, const KArgs& a, double* lds) {
  if constexpr (P::USE) {
    const int lane = threadIdx.x & (WAVE - 1);
    const long long slot = (long long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    double* w = lds + (threadIdx.x / WAVE) * P::WORDS + lane;
    double* ws = a.ws + slot * P::WS_PER_SLOT + lane;
    s.kx = P::KLDS ? w : ws;
    s.zx = w + P::KWORDS;
    s.wx = w + P::KWORDS + P::ZWORDS;
    s.eox = P::ELDS ? w + P::KWORDS + P::ZWORDS + P::WWORDS : ws + P::WS_RATES;
  } else {
    (void)s; (void)a; (void)lds;
  }
}

template <class Net, class M, bool RCP>
using PerCellLane =
    Lane<Cfg<Net, M, RCP, LdsPolicy<Net, M>::KS, false, LdsPolicy<Net, M>::ZX, LdsPolicy<Net, M>::WX>>;

__device__ inline double odv_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline void odv_store_counters(const KArgs& a, long long i, const Counters& c) {
  const long long n = a.ncells;
  a.counters[0 * n + i] = c.attempts;
  a.counters[1 * n + i] = c.steps;
  a.counters[2 * n + i] = c.newton;
  a.counters[3 * n + i] = c.jac;
  a.counters[4 * n + i] = c.errfail;
  a.counters[5 * n + i] = c.convfail;
  a.counters[6 * n + i] = c.stepsel;
  a.counters[7 * n + i] = c.qsum;
}

// y: the integrated species; frozen species keep their value in a.y
template <class L>
__device__ inline void odv_store(const KArgs& a, long long i, const L& s, const double* y) {
  constexpr int N = L::N;
#pragma unroll
  for (int j = 0; j < N; ++j) a.y[(long long)L::Net::ACT[j] * a.ncells + i] = y[j];
  if (a.status) a.status[i] = s.status;
  if (a.dt_out) a.dt_out[i] = s.dt;
  if (a.counters) odv_store_counters(a, i, s.cnt);
}

// Warm restart (no counterpart in the reference, whose SolveODE cold-starts
// every call: src/odevec.F90:170-182). After a successful solve a cell's
// whole integrator state goes to `hist`; the next call on that cell continues the integration
// from it when the call starts where the last one stopped (t0 equal to the
// saved t_stop) on the state it returned (y and parameters bitwise equal to
// the saved ones), else the cell starts cold. Steps are never clamped to
// t_stop (the last step overshoots and the result is interpolated), so a
// solve to t1 continued to t2 takes exactly the steps of one solve to t2.
// The counts the mixed form's FP64 fallback decides on (convergence and
// invariant failures) carry over, and a cell that FP64 lanes continued
// (flag 8) is continued by FP64 lanes again, from its history, so with the
// fallback too a continued solve equals one solve.
template <class Net>
struct Hist {
  static constexpr int N = Net::NACT, NF = Net::NEQ, NP = Net::NPARAM;
  static constexpr int VALID = 0, TSTOP = 1, T = 2, DT = 3, RATE = 4, JCH = 5, CERR = 6;
  // q, steps since Jacobian, lju, failed tests, conv fails, flags, fz, failed invariant tests,
  // convergence failures
  static constexpr int INTS = 7, NINT = 9;
  static constexpr int Y = INTS + NINT;      // the state returned (all NEQ species)
  static constexpr int P = Y + NF;           // the parameters the rates came from
  static constexpr int Z = P + NP;           // Nordsieck history, column-major (c*N + j)
  static constexpr int W = Z + NCOL * N;     // error weights
  static constexpr int EO = W + N;           // previous accepted correction
  static constexpr int A = EO + N;           // P's LU factors
  static constexpr int K = A + Net::NNZ;     // rate coefficients
  static constexpr int T0 = K + Net::NRATE;  // a paused cell (VALID 2): its solve's start time
  static constexpr int CNT = T0 + 1;         // ... and its 8 event counters so far
  static constexpr int PLANES = CNT + 8;
};

// Hand-off record of the mixed form's FP64 fallback (no counterpart in the
// reference): the integrator state in which a cell's convergence failure
// left it when it reached the fallback's limit -- Nordsieck history already
// reset, step size cut, a Jacobian refresh pending -- from which the FP64
// pass continues the solve; cell i at ho[i*REC ...]. The rates and the
// frozen species' flags are recomputed from the cell's input state, which a
// handed-back cell leaves untouched, and the LU factors are not kept (the
// next attempt refreshes them). The event counters travel with the cell.
template <class Net>
struct Handoff {
  static constexpr int N = Net::NACT;
  static constexpr int T = 0, DT = 1, RATE = 2, JCH = 3, CERR = 4;
  static constexpr int INTS = 5, NINT = 7;   // q, swu, lju, fet, fc, flags, failed invariant tests
  static constexpr int CNT = INTS + NINT;    // the 8 event counters
  static constexpr int Z = CNT + 8;          // Nordsieck history, column-major (c*N + j)
  static constexpr int W = Z + NCOL * N;     // error weights
  static constexpr int EO = W + N;           // previous accepted correction
  static constexpr int REC = EO + N;
};

// The record and the list entry pass between wavefronts of one launch that
// may run on different XCDs, whose L2 caches are not coherent with each other:
// they are written and read only with agent-scope atomic accesses (coherent
// without writing back or invalidating the caches, which a release or
// acquire fence would do for the whole L2 at every hand-off), and the entry
// is stored only once the record's stores have completed (odv_finish).
__device__ inline void ho_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ho_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class L>
__device__ inline void odv_handoff_save(const KArgs& a, long long i, const L& s) {
  using H = Handoff<typename L::Net>;
  double* h = a.ho + i * H::REC;
  ho_st(&h[H::T], s.t);
  ho_st(&h[H::DT], s.dt);
  ho_st(&h[H::RATE], s.rate);
  ho_st(&h[H::JCH], s.jchange);
  ho_st(&h[H::CERR], s.cerr);
  ho_st(&h[H::INTS], (double)s.q);
  ho_st(&h[H::INTS + 1], (double)s.swu);
  ho_st(&h[H::INTS + 2], (double)s.lju);
  ho_st(&h[H::INTS + 3], (double)s.fet);
  ho_st(&h[H::INTS + 4], (double)s.fc);
  ho_st(&h[H::INTS + 5], (double)((s.first ? 1 : 0) | (s.upd_jac ? 2 : 0) | (s.neglect_up ? 4 : 0)));
  ho_st(&h[H::INTS + 6], (double)s.cnt.invfail);
  const Counters& c = s.cnt;
  const int cv[8] = {c.attempts, c.steps, c.newton, c.jac, c.errfail, c.convfail, c.stepsel, c.qsum};
#pragma unroll
  for (int k = 0; k < 8; ++k) ho_st(&h[H::CNT + k], (double)cv[k]);
#pragma unroll
  for (int c2 = 0; c2 < NCOL; ++c2)
#pragma unroll
    for (int j = 0; j < H::N; ++j) ho_st(&h[H::Z + c2 * H::N + j], s.z(c2, j));
#pragma unroll
  for (int j = 0; j < H::N; ++j) {
    ho_st(&h[H::W + j], s.w(j));
    ho_st(&h[H::EO + j], s.eo(j));
  }
}

template <class Net>
__device__ inline void odv_handoff_no_counters(const KArgs& a, long long i) {
  using H = Handoff<Net>;
  double* h = a.ho + i * H::REC + H::CNT;
#pragma unroll
  for (int k = 0; k < 8; ++k) ho_st(&h[k], 0.0);
}

// the counters a handed-back cell brought from the launch that stopped it
template <class Net>
__device__ inline Counters odv_handoff_counters(const KArgs& a, long long i) {
  using H = Handoff<Net>;
  const double* h = a.ho + i * H::REC + H::CNT;
  return Counters{(int)ho_ld(h), (int)ho_ld(h + 1), (int)ho_ld(h + 2), (int)ho_ld(h + 3), (int)ho_ld(h + 4),
                  (int)ho_ld(h + 5), (int)ho_ld(h + 6), (int)ho_ld(h + 7), 0};
}

template <class L>
__device__ inline void odv_handoff_resume(const KArgs& a, long long i, L& s, const double* yfull,
                                          const double* prm, double ts) {
  using H = Handoff<typename L::Net>;
  const double* h = a.ho + i * H::REC;
  odv_cell_constants(s, yfull, prm, a.ctl);
  s.t = ho_ld(&h[H::T]);
  s.dt = ho_ld(&h[H::DT]);
  s.rate = ho_ld(&h[H::RATE]);
  s.jchange = ho_ld(&h[H::JCH]);
  s.cerr = ho_ld(&h[H::CERR]);
  odv_set_order(s, (int)ho_ld(&h[H::INTS]));
  s.swu = (int)ho_ld(&h[H::INTS + 1]);
  s.lju = (int)ho_ld(&h[H::INTS + 2]);
  s.fet = (int)ho_ld(&h[H::INTS + 3]);
  s.fc = (int)ho_ld(&h[H::INTS + 4]);
  const int fl = (int)ho_ld(&h[H::INTS + 5]);
  s.first = (fl & 1) != 0;
  s.upd_jac = (fl & 2) != 0;
  s.neglect_up = (fl & 4) != 0;
  s.cnt = odv_handoff_counters<typename L::Net>(a, i);
  s.cnt.invfail = (int)ho_ld(&h[H::INTS + 6]);
  s.status = ST_RUNNING;
  s.tstop = ts;
#pragma unroll
  for (int c = 0; c < NCOL; ++c)
#pragma unroll
    for (int j = 0; j < H::N; ++j) s.set_z(c, j, ho_ld(&h[H::Z + c * H::N + j]));
  bool nan = false;
#pragma unroll
  for (int j = 0; j < H::N; ++j) {
    const double wj = ho_ld(&h[H::W + j]);
    s.set_w(j, wj);
    nan = nan || (wj != wj);
    s.set_eo(j, ho_ld(&h[H::EO + j]));
    s.en[j] = 0.0;
  }
  // weights held in registers carry their NaN flag (as odv_weights sets it)
  if constexpr (!L::WX_) s.fz = (s.fz & ~8) | (nan ? 8 : 0);
}

// cell i's history continues this call: valid, saved where this call starts,
// on the state and parameters it returned
template <class Net>
__device__ inline bool odv_hist_match(const KArgs& a, long long i, const double* y, const double* prm, double t0) {
  using H = Hist<Net>;
  const double* h = a.hist + i * H::PLANES;
  if (h[H::VALID] != 1.0 || h[H::TSTOP] != t0) return false;
  bool same = true;
#pragma unroll
  for (int j = 0; j < H::NF; ++j) same = same && (h[H::Y + j] == y[j]);
#pragma unroll
  for (int p = 0; p < H::NP; ++p) same = same && (h[H::P + p] == prm[p]);
  return same;
}

// FP64 lanes of the fallback continued cell i in the call that saved its history
template <class Net>
__device__ inline bool odv_hist_fp64(const KArgs& a, long long i) {
  using H = Hist<Net>;
  return (((int)a.hist[i * H::PLANES + H::INTS + 5]) & 8) != 0;
}

// the rates saved with cell i's history (computed once, from the state the
// cell's integration started from)
template <class L>
__device__ inline void odv_hist_rates(const KArgs& a, long long i, L& s) {
  using H = Hist<typename L::Net>;
  const double* h = a.hist + i * H::PLANES;
#pragma unroll
  for (int r = 0; r < L::Net::NRATE; ++r) {
    const double v = h[H::K + r];
    if (L::KS == 0) s.kl[r] = v;
    else s.kx[r * L::KS] = v;
  }
}

template <class L>
__device__ inline void odv_hist_resume(const KArgs& a, long long i, L& s, double ts) {
  using H = Hist<typename L::Net>;
  using M = typename L::M;
  const double* h = a.hist + i * H::PLANES;
  s.t = h[H::T];
  s.dt = h[H::DT];
  s.rate = h[H::RATE];
  s.jchange = h[H::JCH];
  s.cerr = h[H::CERR];
  odv_set_order(s, (int)h[H::INTS]);
  s.swu = (int)h[H::INTS + 1];
  s.lju = (int)h[H::INTS + 2];
  s.fet = (int)h[H::INTS + 3];
  s.fc = (int)h[H::INTS + 4];
  const int fl = (int)h[H::INTS + 5];
  s.first = (fl & 1) != 0;
  s.upd_jac = (fl & 2) != 0;
  s.neglect_up = (fl & 4) != 0;
  s.fz = (int)h[H::INTS + 6];
  s.cnt.invfail = (int)h[H::INTS + 7];
  s.cnt.convfail = (int)h[H::INTS + 8];
  s.status = ST_RUNNING;
  s.tstop = ts;
#pragma unroll
  for (int c = 0; c < NCOL; ++c)
#pragma unroll
    for (int j = 0; j < H::N; ++j) s.set_z(c, j, h[H::Z + c * H::N + j]);
#pragma unroll
  for (int j = 0; j < H::N; ++j) {
    s.set_w(j, h[H::W + j]);
    s.set_eo(j, h[H::EO + j]);
    s.en[j] = 0.0;
  }
#pragma unroll
  for (int e = 0; e < L::Net::NNZ; ++e) s.a[e] = (M)h[H::A + e];
  odv_hist_rates(a, i, s);
}

// after odv_store: a.y holds the state returned to the caller; fp64: the
// cell ran on an FP64 lane of the fallback
template <class L>
__device__ inline void odv_hist_write(const KArgs& a, long long i, const L& s, bool fp64);

template <class L>
__device__ inline void odv_hist_save(const KArgs& a, long long i, const L& s, bool fp64) {
  using H = Hist<typename L::Net>;
  double* h = a.hist + i * H::PLANES;
  if (s.status == ST_SKIPPED) return;   // nothing integrated: the saved state stays usable
  if (s.status != ST_OK) {
    h[H::VALID] = 0.0;
    return;
  }
  odv_hist_write(a, i, s, fp64);
}

// the cell's whole integrator state, the state in a.y and its parameters
template <class L>
__device__ inline void odv_hist_write(const KArgs& a, long long i, const L& s, bool fp64) {
  using H = Hist<typename L::Net>;
  const long long n = a.ncells;
  double* h = a.hist + i * H::PLANES;
  h[H::TSTOP] = s.tstop;
  h[H::T] = s.t;
  h[H::DT] = s.dt;
  h[H::RATE] = s.rate;
  h[H::JCH] = s.jchange;
  h[H::CERR] = s.cerr;
  h[H::INTS] = (double)s.q;
  h[H::INTS + 1] = (double)s.swu;
  h[H::INTS + 2] = (double)s.lju;
  h[H::INTS + 3] = (double)s.fet;
  h[H::INTS + 4] = (double)s.fc;
  h[H::INTS + 5] = (double)((s.first ? 1 : 0) | (s.upd_jac ? 2 : 0) | (s.neglect_up ? 4 : 0) | (fp64 ? 8 : 0));
  h[H::INTS + 6] = (double)s.fz;
  h[H::INTS + 7] = (double)s.cnt.invfail;
  h[H::INTS + 8] = (double)s.cnt.convfail;
#pragma unroll
  for (int j = 0; j < H::NF; ++j) h[H::Y + j] = a.y[(long long)j * n + i];
#pragma unroll
  for (int p = 0; p < H::NP; ++p) h[H::P + p] = a.prm[p * a.ps_plane + i * a.ps_cell];
#pragma unroll
  for (int c = 0; c < NCOL; ++c)
#pragma unroll
    for (int j = 0; j < H::N; ++j) h[H::Z + c * H::N + j] = s.z(c, j);
#pragma unroll
  for (int j = 0; j < H::N; ++j) {
    h[H::W + j] = s.w(j);
    h[H::EO + j] = s.eo(j);
  }
#pragma unroll
  for (int e = 0; e < L::Net::NNZ; ++e) h[H::A + e] = (double)s.a[e];
#pragma unroll
  for (int r = 0; r < L::Net::NRATE; ++r) h[H::K + r] = s.kptr()[r * (L::KS > 0 ? L::KS : 1)];
  h[H::VALID] = 1.0;
}

// A cell whose launch quantum ran out: its whole state (the input state
// and parameters it started from, which stay in place, and its counters so
// far) into its history as a pause record (VALID 2), which the next launch
// of the same solve continues.
template <class L>
__device__ inline void odv_hist_pause(const KArgs& a, long long i, const L& s) {
  using H = Hist<typename L::Net>;
  odv_hist_write(a, i, s, false);
  double* h = a.hist + i * H::PLANES;
  h[H::T0] = a.t0[i * a.t0_s];
  const Counters& c = s.cnt;
  const int cv[8] = {c.attempts, c.steps, c.newton, c.jac, c.errfail, c.convfail, c.stepsel, c.qsum};
#pragma unroll
  for (int k = 0; k < 8; ++k) h[H::CNT + k] = (double)cv[k];
  h[H::VALID] = 2.0;
}

// the pause record of cell i continues this launch (the same solve: start
// and end time, state and parameters unchanged)
template <class L>
__device__ inline bool odv_hist_unpause(const KArgs& a, long long i, L& s, const double* y, const double* prm,
                                        double t0, double ts) {
  using H = Hist<typename L::Net>;
  const double* h = a.hist + i * H::PLANES;
  if (h[H::VALID] != 2.0 || h[H::T0] != t0 || h[H::TSTOP] != ts) return false;
  bool same = true;
#pragma unroll
  for (int j = 0; j < H::NF; ++j) same = same && (h[H::Y + j] == y[j]);
#pragma unroll
  for (int p = 0; p < H::NP; ++p) same = same && (h[H::P + p] == prm[p]);
  if (!same) return false;
  odv_hist_resume(a, i, s, ts);
  s.cnt = Counters{(int)h[H::CNT], (int)h[H::CNT + 1], (int)h[H::CNT + 2], (int)h[H::CNT + 3], (int)h[H::CNT + 4],
                   (int)h[H::CNT + 5], (int)h[H::CNT + 6], (int)h[H::CNT + 7], (int)h[H::INTS + 7]};
  return true;
}

// Load cell i into the lane and start its solve; returns true if there is
// nothing to integrate (masked out or t_stop <= t0; y is then left as is).
// HIST: the kernel variant that reads and writes a warm-restart history (a
// separate instantiation: the history code raises the register pressure of
// the refill path, which the kernels without it should not pay); 2: also
// the launch quantum's pause and resume (the batch kernel's third variant:
// that code more than tripled the history variant's spills, 552 -> 1,928 B
// per lane for the MIX3 bundle).
template <class L, int HIST = 0>
__device__ inline bool odv_load_begin(const KArgs& a, long long i, L& s) {
  using Net = typename L::Net;
  constexpr int NF = Net::NEQ;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  double y[NF], prm[NP];
#pragma unroll
  for (int j = 0; j < NF; ++j) y[j] = a.y[(long long)j * a.ncells + i];
#pragma unroll
  for (int p = 0; p < NP; ++p) prm[p] = (Net::NPARAM > 0) ? a.prm[p * a.ps_plane + i * a.ps_cell] : 0.0;
  const double t0 = a.t0[i * a.t0_s];
  const double ts = a.tstop[i * a.ts_s];
  const double dh = a.dt0[i * a.dt_s];
  s.active = true;
  s.cnt = Counters{0, 0, 0, 0, 0, 0, 0, 0};
  s.dt = dh;
  if (a.times) a.times[i] = (long long)wall_clock64();
  if constexpr (HIST == 2) {
    if (a.quantum > 0 && a.hist && a.status) {
      // a solve in quantum launches: the cell's phase is its status entry
      const int ph = a.status[i];
      if (ph == ST_PAUSED && odv_hist_unpause(a, i, s, y, prm, t0, ts)) return false;
      if (ph != ST_FRESH && ph != ST_PAUSED) {
        s.status = ST_DONE_EARLIER;
        return true;
      }
    }
  }
  if ((a.mask && !a.mask[i]) || !(t0 < ts)) {
    s.status = ST_SKIPPED;
    return true;
  }
  if constexpr (HIST) {
    // (in quantum launches the history holds pause records only: a fresh
    // cell starts cold)
    if (a.hist && a.quantum == 0 && odv_hist_match<Net>(a, i, y, prm, t0)) {
      const bool f64 = odv_hist_fp64<Net>(a, i);
      if (a.ho_mode == 1 && f64) {
        // FP64 lanes continued the cell last call: they continue it again
        s.status = ST_PROMOTE_HIST;
        return true;
      }
      // (an FP64 lane of the fallback takes a cell handed back in this call
      // from its hand-off record, below)
      if (a.ho_mode != 2 || f64) {
        odv_hist_resume(a, i, s, ts);
        if (s.t < s.tstop) return false;
        s.status = ST_OK;   // the last step already passed the new t_stop: interpolate
        return true;
      }
    }
  }
  if constexpr (sizeof(typename L::M) == sizeof(double)) {
    // the FP64 fallback's pass: the cell continues where the launch that
    // handed it back stopped it
    if (a.ho_mode == 2 && a.ho) {
      odv_handoff_resume(a, i, s, y, prm, ts);
      if constexpr (HIST) {
        // a cell this call continued from its history before handing it
        // back: its rates are the history's (the rates come from the state
        // the integration started from, not from this call's input state)
        if (a.hist && a.quantum == 0 && odv_hist_match<Net>(a, i, y, prm, t0)) odv_hist_rates(a, i, s);
      }
      return false;
    }
  }
  LaneGroup g;
  odv_begin(s, y, prm, t0, ts, dh, a.ctl, g);
  return false;
}

template <class L, int HIST = 0>
__device__ inline void odv_finish(const KArgs& a, long long i, const L& s) {
  if constexpr (HIST == 2) {
    if (s.status == ST_DONE_EARLIER) return;
    if (s.status == ST_PAUSED) {
      odv_hist_pause(a, i, s);
      if (a.status) a.status[i] = ST_PAUSED;
      return;
    }
  }
  if (a.times) a.times[a.ncells + i] = (long long)wall_clock64();
  if (s.status == ST_PROMOTE || (HIST && s.status == ST_PROMOTE_HIST)) {
    // the caller's state stays untouched for the FP64 pass over promo[],
    // which continues the cell from its hand-off record (or, ST_PROMOTE_HIST,
    // from its history; the record then only says that no work came with it)
    if (a.ho) {
      if constexpr (HIST) {
        if (s.status == ST_PROMOTE) odv_handoff_save(a, i, s);
        else odv_handoff_no_counters<typename L::Net>(a, i);
      } else {
        odv_handoff_save(a, i, s);
      }
    }
    // the entry is published once the record's stores have completed (an
    // FP64 lane of the same launch may be waiting for it)
    __builtin_amdgcn_s_waitcnt(0);
    const unsigned k = atomicAdd(a.promo_n, 1u);
    __hip_atomic_store(&a.promo[k], (int)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (s.status == ST_SKIPPED) {
    if (a.status) a.status[i] = s.status;
    if (a.dt_out) a.dt_out[i] = s.dt;
    if (a.counters) odv_store_counters(a, i, s.cnt);
    return;
  }
  double y[L::N];
  odv_result(s, y);
  odv_store(a, i, s, y);
  if constexpr (HIST) {
    if (a.hist && a.quantum == 0) odv_hist_save(a, i, s, a.ho_mode == 2);
  }
}

template <class Net, class M, bool RCP, int HIST = 0>
__global__ void __launch_bounds__(MAX_BLOCK) odv_k_percell(KArgs a) {
  using P = LdsPolicy<Net, M>;
  using L = PerCellLane<Net, M, RCP>;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double fl64 = 0.0, fl32 = 0.0;
  if (i < a.ncells) {
    L s;
    odv_bind_columns<P>(s, a, odv_lane_lds);
    LaneGroup g;
    bool done = odv_load_begin<L, HIST>(a, i, s);
    while (!done) done = odv_attempt(s, a.ctl, g);
    odv_finish<L, HIST>(a, i, s);
    if (a.flops && s.status != ST_SKIPPED) odv_flops<L>(s.cnt, fl64, fl32);
  }
  if (a.flops) {
    fl64 = odv_wave_sum(fl64);
    fl32 = odv_wave_sum(fl32);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&a.flops[0], fl64);
      atomicAdd(&a.flops[1], fl32);
    }
  }
}

// Persistent form. Every lane starts on one queue position; the queue
// counter starts at the number of launched lanes. Whenever at least one lane
// of a wavefront has finished, the wavefront reserves that many positions
// with a single atomic and hands them out in lane order. Position p holds
// cell a.order[p] (costliest first) or cell p. With an order, lane l of
// wavefront w starts on position l*W + w (W wavefronts): the costliest cells
// start at once, one per wavefront, rather than together in the first
// wavefronts (cells that run side by side in one wavefront run at the pace
// of the union of their phases). Every lane's loop ends: each iteration
// either advances a cell by one attempt (cells are bounded by max_attempts)
// or, once the wavefront has seen the queue exhausted and holds no cell,
// leaves. `lds` holds the workgroup's wavefront slices,
// LdsPolicy<Net, M>::WORDS doubles each.
// (forced inline: called from more than one kernel of a unit, the compiler
// would otherwise emit it as a function, with the arguments in scratch memory
// and a call per launch)
template <class Net, class M, bool RCP, int HIST = 0>
__device__ __forceinline__ void odv_persistent_run(const KArgs& a, double* lds, long long nl = 0) {
  using P = LdsPolicy<Net, M>;
  using L = PerCellLane<Net, M, RCP>;
  // lanes taking part (the first nl of the grid; 0: all)
  const long long nlanes = nl > 0 ? nl : (long long)gridDim.x * blockDim.x;
  const long long gl = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long nq = a.qlen ? (long long)*a.qlen : a.ncells;   // queue positions
  const long long p0 = (a.order || a.spread) ? (long long)lane * (nlanes >> 6) + (gl >> 6) : gl;
  long long i = p0 < nq ? (a.order ? (long long)a.order[p0] : p0) : a.ncells;
  // a list entry below 0: a cell already taken elsewhere (the fallback's list)
  L s;
  odv_bind_columns<P>(s, a, lds);
  LaneGroup g;
  double fl64 = 0.0, fl32 = 0.0;
  bool have = p0 < nq && i >= 0;
  bool done = true;
  bool exhausted = nlanes >= nq;   // wave-uniform: no cell left to hand out
  int left = a.quantum;            // attempts left in this launch for the lane's cell (quantum launches)
  if (have) done = odv_load_begin<L, HIST>(a, i, s);
  while (true) {
    const bool need = !have || done;
    if (have && done) {
      odv_finish<L, HIST>(a, i, s);
      if (a.flops && s.status != ST_SKIPPED) {
        double f64, f32;
        Counters c = s.cnt;
        if (a.ho_mode == 2 && a.ho) {
          // a continued cell: count the work of this launch only
          const Counters b = odv_handoff_counters<Net>(a, i);
          c = Counters{c.attempts - b.attempts, c.steps - b.steps, c.newton - b.newton, c.jac - b.jac,
                       c.errfail - b.errfail, c.convfail - b.convfail, c.stepsel - b.stepsel, c.qsum - b.qsum, 0};
        }
        odv_flops<L>(c, f64, f32);
        fl64 += f64;
        fl32 += f32;
      }
      have = false;
    }
    const unsigned long long ball = __ballot(need);
    // refill once REFILL_MIN lanes wait, or when every lane does: the queue
    // atomic and the wavefront's pass through the new cells' start (rates,
    // initial step) then serve several lanes at once
    if (ball && !exhausted && (__popcll(ball) >= REFILL_MIN || ball == __ballot(true))) {
      const int cnt = __popcll(ball);
      const int rank = __popcll(ball & ((1ull << lane) - 1ull));
      const int leader = __ffsll((long long)ball) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)cnt);
      base = __shfl(base, leader, 64);
      if ((long long)base + cnt >= nq) exhausted = true;
      if (need) {
        const long long ni = (long long)base + rank;
        if (ni < nq) {
          i = a.order ? (long long)a.order[ni] : ni;
          if (i >= 0) {
            have = true;
            left = a.quantum;
            done = odv_load_begin<L, HIST>(a, i, s);
          }
        }
      }
    }
    if (__ballot(have) == 0ull) break;
    if (have && !done) {
      done = odv_attempt(s, a.ctl, g);
      if constexpr (HIST == 2) {
        if (!done && a.quantum > 0 && --left <= 0) {
          s.status = ST_PAUSED;   // the quantum is used up: the next launch continues the cell
          done = true;
        }
      }
    }
  }
  if (a.flops) {
    fl64 = odv_wave_sum(fl64);
    fl32 = odv_wave_sum(fl32);
    if (lane == 0) {
      atomicAdd(&a.flops[0], fl64);
      atomicAdd(&a.flops[1], fl32);
    }
  }
  if (a.fb_done) {
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) __hip_atomic_fetch_add(a.fb_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The FP64 fallback beside the launch: cells handed back by the wavefronts
// of a persistent loop running at the same time (a kernel on another stream;
// their list b.order, filled in hand-back order, entries -1 until written;
// count *b.qlen) are continued from their hand-off records by FP64 lanes
// while that loop still runs. Each lane
// claims the next list position (b.queue, one atomic per wavefront), waits
// until the position's entry is written, marks the entry taken (-2) and
// runs the cell. A claim past the final count (known once all `nprod`
// mixed wavefronts have counted themselves in *done) is void; a wavefront
// with nothing to run sleeps between looks at the list, and one that has
// waited for `idle_rounds` such looks in a row drops its pending claims and
// leaves -- the entries it did not take stay in the list for a pass after
// the launch, so a wait can never lose a cell. The wait is bounded by a
// count, not by the device clock: with the clock read (s_memrealtime) in
// this loop, this compiler built a consumer kernel whose continued cells
// differed from the same loop compiled as a called function (a wrong result
// in the build without contraction: 29 of 113 PRIMORDIAL cells, the same
// ones at every run); without the read they agree bit for bit. Columns: b.ws
// (a workspace for the FP64 layout) and the wavefront's LDS slice.
template <class Net, int HIST = 0>
__device__ __forceinline__ void odv_fb_consume(const KArgs& b, double* lds, const unsigned* done, unsigned nprod,
                                               long long idle_rounds) {
  using P = LdsPolicy<Net, double>;
  using L = PerCellLane<Net, double, false>;
  const int lane = threadIdx.x & 63;
  int* list = const_cast<int*>(b.order);
  L s;
  odv_bind_columns<P>(s, b, lds);
  LaneGroup g;
  double fl64 = 0.0, fl32 = 0.0;
  long long i = 0;
  bool have = false, fin = true;
  long long pos = -1;      // claimed list position; -1 none, -2 void
  bool closed = false;     // wave-uniform: claim no more
  long long idle = 0;       // looks at the list in a row with nothing to run
  while (true) {
    if (have && fin) {
      odv_finish<L, HIST>(b, i, s);
      if (b.flops && s.status != ST_SKIPPED) {
        const Counters c0 = odv_handoff_counters<Net>(b, i);
        const Counters& c = s.cnt;
        const Counters d{c.attempts - c0.attempts, c.steps - c0.steps, c.newton - c0.newton, c.jac - c0.jac,
                         c.errfail - c0.errfail, c.convfail - c0.convfail, c.stepsel - c0.stepsel,
                         c.qsum - c0.qsum, 0};
        double f64, f32;
        odv_flops<L>(d, f64, f32);
        fl64 += f64;
        fl32 += f32;
      }
      have = false;
    }
    const bool need = !have && pos == -1;
    const unsigned long long ball = __ballot(need);
    if (ball && !closed) {
      const int cnt = __popcll(ball);
      const int rank = __popcll(ball & ((1ull << lane) - 1ull));
      const int leader = __ffsll((long long)ball) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(b.queue, (unsigned long long)cnt);
      base = __shfl(base, leader, 64);
      if (need) pos = (long long)base + rank < b.ncells ? (long long)base + rank : -2;
    }
    // look at the claimed entries (the record is read, by address from the
    // entry, only once the entry is there)
    int v = -1;
    bool all_in = false;
    if (pos >= 0) {
      v = __hip_atomic_load(&list[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (v < 0) all_in = __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nprod;
    }
    if (pos >= 0) {
      if (v < 0 && all_in) {
        // every mixed wavefront has finished: an entry below the final count
        // is written by now, one at or past it never will be (void claim)
        v = __hip_atomic_load(&list[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v < 0 && pos >= (long long)__hip_atomic_load(b.qlen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
          pos = -2;
      }
      if (v >= 0) {
        __hip_atomic_store(&list[pos], -2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i = v;
        have = true;
        fin = odv_load_begin<L, HIST>(b, i, s);
        pos = -1;
      }
    }
    // claims are handed out in order: once one is void, all later ones are
    if (__ballot(pos == -2)) closed = true;
    if (pos == -2) pos = -3;   // settled
    const bool running = have && !fin;
    if (__ballot(running)) {
      idle = 0;
      if (running) fin = odv_attempt(s, b.ctl, g);
    } else {
      if (closed && __ballot(pos >= 0) == 0ull && __ballot(have) == 0ull) break;
      if (++idle > idle_rounds) {
        closed = true;    // leave the pending claims to the pass after the launch
        if (pos >= 0) pos = -3;
      }
      __builtin_amdgcn_s_sleep(4);
    }
  }
  if (b.flops) {
    fl64 = odv_wave_sum(fl64);
    fl32 = odv_wave_sum(fl32);
    if (lane == 0) {
      atomicAdd(&b.flops[0], fl64);
      atomicAdd(&b.flops[1], fl32);
    }
  }
}

// Mixed (or FP64) persistent loop with its FP64 fallback in the same launch:
// the first nprod wavefronts run the persistent loop over a's queue, handing
// cells back (a.ctl.promote_cf, a.ho_mode 1), then join the last wavefronts,
// which continue handed-back cells from the start (odv_fb_consume, args b).
// One wavefront per workgroup: both layouts use the slice at the LDS base.
// The FP64 lanes of the fallback as a kernel of their own, launched on a
// second stream beside the persistent kernel whose hand-backs they take
// (nprod: that kernel's wavefronts); one wavefront per workgroup.
template <class Net, int HIST = 0>
__global__ void __launch_bounds__(WAVE) odv_k_fb_consume(KArgs b, unsigned* done, unsigned nprod,
                                                         long long idle_rounds) {
  odv_fb_consume<Net, HIST>(b, odv_lane_lds, done, nprod, idle_rounds);
}

// With a.order and a.qlen the queue is the first *a.qlen entries of a.order
// (the FP64 pass over the mixed form's promoted cells).
template <class Net, class M, bool RCP, int HIST = 0>
__global__ void __launch_bounds__(MAX_BLOCK) odv_k_persistent(KArgs a) {
  odv_persistent_run<Net, M, RCP, HIST>(a, odv_lane_lds);
}

// ------------------------------------------------------- network functions
// The generated plugin's GetRHS and GetJac (src/odevec_commons.F90:32-95)
// for a block of cells, in the reference's layouts: rhs[j*ncells + i] over
// all NEQ species (frozen ones 0); jac[e*ncells + i] over the full system's
// GetJac storage (Net::JMAP). The rates come from each cell's own state and
// parameters, as in the solver.
template <class Net>
__global__ void k_get_rhs(KArgs a, double* rhs) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncells) return;
  constexpr int NF = Net::NEQ, N = Net::NACT;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  double y[NF], prm[NP], k[Net::NRATE], ya[N], f[N], out[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) y[j] = a.y[(long long)j * a.ncells + i];
#pragma unroll
  for (int p = 0; p < NP; ++p) prm[p] = (Net::NPARAM > 0) ? a.prm[p * a.ps_plane + i * a.ps_cell] : 0.0;
  Net::rates(y, prm, k);
#pragma unroll
  for (int j = 0; j < N; ++j) ya[j] = y[Net::ACT[j]];
  Net::template rhs<1>(ya, k, f);
#pragma unroll
  for (int j = 0; j < NF; ++j) out[j] = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j) out[Net::ACT[j]] = f[j];
#pragma unroll
  for (int j = 0; j < NF; ++j) rhs[(long long)j * a.ncells + i] = out[j];
}

template <class Net>
__global__ void k_get_jac(KArgs a, double hb, double* jac) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ncells) return;
  constexpr int NF = Net::NEQ, N = Net::NACT;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  double y[NF], prm[NP], k[Net::NRATE], ya[N], P[Net::NNZ];
#pragma unroll
  for (int j = 0; j < NF; ++j) y[j] = a.y[(long long)j * a.ncells + i];
#pragma unroll
  for (int p = 0; p < NP; ++p) prm[p] = (Net::NPARAM > 0) ? a.prm[p * a.ps_plane + i * a.ps_cell] : 0.0;
  Net::rates(y, prm, k);
#pragma unroll
  for (int j = 0; j < N; ++j) ya[j] = y[Net::ACT[j]];
  Net::template jac<double, 1>(ya, k, hb, P);
  for (int e = 0; e < Net::JMAP_N; ++e) {
    const int m = Net::JMAP[e];
    jac[(long long)e * a.ncells + i] = m >= 0 ? P[m] : (m == -2 ? 1.0 : 0.0);
  }
}

// ------------------------------------------------------------- cost order
// Persistent-mode scheduling. A launch lasts at least as long as its slowest
// cell (PRIMORDIAL, 1M cells: one cell takes 41,008 step attempts, ~320 ms
// alone, the median 264), so the queue hands out the costliest cells first.
// Proxy for a cell's cost: log2((t_stop - t0)/h0), h0 the reference's
// initial step size (CalcInitialStepSize, src/odevec.F90:216-238): a stiff
// cell over a long span needs many steps (correlation 0.89 with the log of
// the attempts on that workload). The cells are binned by the proxy (2 bins
// per octave over 0..127 octaves, costliest bin first; that workload's
// proxies span 10..128 octaves, and with 4 bins per octave, saturating at 63,
// half of its cells -- among them every cell the mixed form hands to its
// FP64 fallback -- shared the top bin in arbitrary order) and
// counting-sorted into `order`; within a bin the order is arbitrary. Only
// the order in which cells start changes: every cell's arithmetic is the
// same.
constexpr int COST_BINS = 256;

template <int UNUSED = 0>
__global__ void k_cost_zero(unsigned* count) {
  if (threadIdx.x < COST_BINS) count[threadIdx.x] = 0u;
}

// The bin counts and the scatter's positions are gathered per workgroup in
// LDS first, so that each workgroup adds to a global counter once per bin it
// holds instead of once per cell (one global atomic per cell on 256 counters
// serialised the two kernels: 6.5 ms each for 1M cells).
template <class Net>
__global__ void k_cost_bin(KArgs a, unsigned char* bin, unsigned* count) {
  __shared__ unsigned h[COST_BINS];
  for (int t = threadIdx.x; t < COST_BINS; t += blockDim.x) h[t] = 0u;
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int NF = Net::NEQ, N = Net::NACT;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  const bool in = i < a.ncells;
  const double t0 = in ? a.t0[i * a.t0_s] : 0.0, ts = in ? a.tstop[i * a.ts_s] : 0.0;
  int b = COST_BINS - 1;   // skipped cells last
  if (in && !(a.mask && !a.mask[i]) && t0 < ts) {
    double y[NF], prm[NP], k[Net::NRATE], ya[N], f[N];
#pragma unroll
    for (int j = 0; j < NF; ++j) y[j] = a.y[(long long)j * a.ncells + i];
#pragma unroll
    for (int p = 0; p < NP; ++p) prm[p] = (Net::NPARAM > 0) ? a.prm[p * a.ps_plane + i * a.ps_cell] : 0.0;
    Net::rates(y, prm, k);
#pragma unroll
    for (int j = 0; j < N; ++j) ya[j] = y[Net::ACT[j]];
    Net::template rhs<1>(ya, k, f);
    const double tol = a.ctl.rtol;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double v = 1.0 / (tol * fabs(ya[j]) + a.ctl.atol);
      acc += f[j] * f[j] * (v * v) * tol * tol;
    }
    const double W0 = fmax(fabs(t0), fabs(ts));
    const double h0 = fmin(sqrt(tol / (1.0 / (W0 * W0) + acc / NF)), ts - t0);
    const double q = log2((ts - t0) / h0);
    const int qb = (q == q) ? (int)fmin(fmax(q * 2.0, 0.0), (double)(COST_BINS - 2)) : 0;
    b = COST_BINS - 2 - qb;
  }
  if (in) {
    bin[i] = (unsigned char)b;
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < COST_BINS; t += blockDim.x)
    if (h[t]) atomicAdd(&count[t], h[t]);
}

// exclusive prefix sum of the bin counts (one workgroup of COST_BINS lanes)
template <int UNUSED = 0>
__global__ void __launch_bounds__(COST_BINS) k_cost_scan(unsigned* count) {
  __shared__ unsigned s[COST_BINS];
  const int t = threadIdx.x;
  s[t] = count[t];
  __syncthreads();
  for (int off = 1; off < COST_BINS; off <<= 1) {
    const unsigned v = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  count[t] = s[t] - count[t];
}

// each workgroup reserves one range per bin (its cells' count) and places
// its cells in it; offset[] starts as the exclusive scan of the bin counts
template <int UNUSED = 0>
__global__ void k_cost_scatter(long long ncells, const unsigned char* bin, unsigned* offset, int* order) {
  __shared__ unsigned h[COST_BINS], base[COST_BINS];
  for (int t = threadIdx.x; t < COST_BINS; t += blockDim.x) h[t] = 0u;
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < ncells;
  const int b = in ? (int)bin[i] : 0;
  const unsigned local = in ? atomicAdd(&h[b], 1u) : 0u;
  __syncthreads();
  for (int t = threadIdx.x; t < COST_BINS; t += blockDim.x)
    if (h[t]) base[t] = atomicAdd(&offset[t], h[t]);
  __syncthreads();
  if (in) order[base[b] + local] = (int)i;
}

// position of each cell in the queue: inv[order[p]] = p
template <int UNUSED = 0>
__global__ void k_invert(long long n, const int* order, int* inv) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) inv[order[p]] = (int)p;
}

// The cost-ordered queue's cells copied into queue order: position p of
// the copies holds cell order[p], so the cells a wavefront takes at one
// refill (consecutive positions) are consecutive in memory, and the
// persistent kernel's state loads and stores are coalesced. The permutation
// happens here once, in streaming passes that read coalesced and write
// scattered (partial cache lines of scattered writes merge in the caches;
// scattered reads would fetch a whole line per value). Broadcast inputs
// (stride 0) and absent outputs are left out (null destination).
template <int UNUSED = 0>
__global__ void k_gather(KArgs a, int neq, int np, const int* inv, double* gy, double* gts, double* gdt,
                         double* gt0, double* gprm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = a.ncells;
  if (i >= n) return;
  const long long p = inv[i];
  for (int j = 0; j < neq; ++j) gy[(long long)j * n + p] = a.y[(long long)j * n + i];
  if (gts) gts[p] = a.tstop[i * a.ts_s];
  if (gdt) gdt[p] = a.dt0[i * a.dt_s];
  if (gt0) gt0[p] = a.t0[i * a.t0_s];
  if (gprm)
    for (int q = 0; q < np; ++q) gprm[(long long)q * n + p] = a.prm[q * a.ps_plane + i * a.ps_cell];
}

// the results back to the caller's cell order
template <int UNUSED = 0>
__global__ void k_scatter(long long n, int neq, const int* order, const double* gy, const int* gst,
                          const double* gdto, const int* gcnt, double* y, int* status, double* dt_out,
                          int* counters) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const long long i = order[p];
  for (int j = 0; j < neq; ++j) y[(long long)j * n + i] = gy[(long long)j * n + p];
  if (status) status[i] = gst[p];
  if (dt_out) dt_out[i] = gdto[p];
  if (counters)
    for (int m = 0; m < 8; ++m) counters[(long long)m * n + i] = gcnt[(long long)m * n + p];
}

// ---------------------------------------------------------------- batches
// One launch over a table of segments, each a block of cells of one network
// in one precision (its own arrays, tolerances and queue counter). Every
// wavefront works through the segments in table order: for each it runs the
// persistent loop on that segment's queue, then moves to the next. The
// segment kind selects one of the bundle's compiled (network, precision)
// entries; the kind is uniform over the workgroup, so the wavefront executes
// one network's code at a time. Segments ordered by decreasing cost per cell
// (the host sorts them) leave cheap cells for the end of the launch.
template <class Net_, class M_, bool RCP_>
struct Entry {
  using Net = Net_;
  using M = M_;
  static constexpr bool RCP = RCP_;
  static constexpr int WORDS = LdsPolicy<Net_, M_>::WORDS;
  static constexpr long long WS_PER_SLOT = LdsPolicy<Net_, M_>::WS_PER_SLOT;
};

struct SegDev {
  KArgs a;      // the segment's cells (queue: its own counter)
  int kind;     // index into the bundle's entries
  int pad;
};

template <class... E>
struct MaxWords;
template <class E>
struct MaxWords<E> {
  static constexpr int WORDS = E::WORDS;
  static constexpr long long WS = E::WS_PER_SLOT;
};
template <class E, class F, class... R>
struct MaxWords<E, F, R...> {
  static constexpr int WORDS = E::WORDS > MaxWords<F, R...>::WORDS ? E::WORDS : MaxWords<F, R...>::WORDS;
  static constexpr long long WS = E::WS_PER_SLOT > MaxWords<F, R...>::WS ? E::WS_PER_SLOT : MaxWords<F, R...>::WS;
};

template <int K, int HIST, class E, class... R>
__device__ inline void odv_batch_dispatch(int kind, const KArgs& a, double* lds) {
  if (kind == K) {
    odv_persistent_run<typename E::Net, typename E::M, E::RCP, HIST>(a, lds);
  } else {
    if constexpr (sizeof...(R) > 0) odv_batch_dispatch<K + 1, HIST, R...>(kind, a, lds);
  }
}

// HIST: the variant for tables where segments carry warm-restart histories
// (segments without one start cold in it too); 2: some segment runs with a
// launch quantum
template <int HIST, class... E>
__global__ void __launch_bounds__(MAX_BLOCK) odv_k_batch(const SegDev* segs, const int* nseg) {
  const int n = *nseg;
  for (int k = 0; k < n; ++k) {
    const KArgs a = segs[k].a;
    // the entry carves its own wavefront slices (at most MaxWords doubles each)
    odv_batch_dispatch<0, HIST, E...>(segs[k].kind, a, odv_lane_lds);
    __syncthreads();   // the next segment's kind may carve LDS differently
  }
}

// Queue counters of a batch: each segment's starts at the launched lane count
// (a kernel, not a memset, so that the launch can be captured in a graph).
template <int UNUSED = 0>
__global__ void k_batch_reset(SegDev* segs, const int* nseg, unsigned long long nlanes) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < *nseg) *segs[k].a.queue = nlanes;
}

// Lockstep group reductions over one workgroup (group <= blockDim.x <= 256).
struct BlockGroup {
  double* red;   // LDS scratch [blockDim/64]
  int* redi;
  bool valid;
  __device__ double max_active(double v, bool act) const {
    double x = act ? v : -INFINITY;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = x;
    __syncthreads();
    double r = red[0];
    for (int k = 1; k < nw; ++k) r = fmax(r, red[k]);
    return r;
  }
  __device__ bool any_of(bool b) const {
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const bool wb = __ballot(b) != 0ull;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) redi[w] = wb ? 1 : 0;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < nw; ++k) r |= redi[k];
    return r != 0;
  }
  __device__ bool any_valid(bool b) const { return any_of(b && valid); }
  __device__ bool any_active(bool b) const { return any_of(b); }
  __device__ double species_sum(double v) const { return v; }
  __device__ double species_max(double v) const { return v; }
};

// Lockstep form. Every decision is taken from group-reduced values, so all
// lanes of the workgroup follow one control path and reach the same barriers.
template <class Net, class M, bool RCP>
__global__ void __launch_bounds__(256) odv_k_lockstep(KArgs a) {
  using L = Lane<Cfg<Net, M, RCP, 0, true>>;
  __shared__ double red[4];
  __shared__ int redi[4];
  constexpr int N = Net::NEQ;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  const long long g0 = (long long)blockIdx.x * a.group;
  const long long i = g0 + threadIdx.x;
  const bool valid = threadIdx.x < a.group && i < a.ncells;
  BlockGroup g{red, redi, valid};
  L s;
  double y[N];
  double prm[NP];
  const long long ii = valid ? i : g0;   // padding lanes mirror the group's first cell
#pragma unroll
  for (int j = 0; j < N; ++j) y[j] = a.y[(long long)j * a.ncells + ii];
#pragma unroll
  for (int p = 0; p < NP; ++p) prm[p] = (Net::NPARAM > 0) ? a.prm[p * a.ps_plane + ii * a.ps_cell] : 0.0;
  // the group shares time, t_stop and the dt hint: those of its first cell
  const double t0 = a.t0[g0 * a.t0_s];
  const double ts = a.tstop[g0 * a.ts_s];
  const double dh = a.dt0[g0 * a.dt_s];
  s.active = valid && (!a.mask || a.mask[ii]);
  s.cnt = Counters{0, 0, 0, 0, 0, 0, 0, 0};
  s.dt = dh;
  s.status = ST_SKIPPED;
  const bool any_act = g.any_of(s.active);
  if (any_act && t0 < ts) {
    odv_begin(s, y, prm, t0, ts, dh, a.ctl, g);
    bool done = false;
    while (!done) done = odv_attempt(s, a.ctl, g);
    if (s.active) {
      double ya[L::N];
      odv_result(s, ya);
#pragma unroll
      for (int j = 0; j < L::N; ++j) y[Net::ACT[j]] = ya[j];
    } else {
      s.status = ST_SKIPPED;
    }
  }
  if (valid) {
    if (s.active) {
#pragma unroll
      for (int j = 0; j < N; ++j) a.y[(long long)j * a.ncells + i] = y[j];
    }
    if (a.status) a.status[i] = s.status;
    if (a.dt_out) a.dt_out[i] = s.dt;
    if (a.counters) odv_store_counters(a, i, s.cnt);
  }
  if (a.flops && threadIdx.x == 0 && s.status != ST_SKIPPED) {
    // every lane of the group does the group's work
    const double nval = (double)min((long long)a.group, a.ncells - g0);
    double f64, f32;
    odv_flops<L>(s.cnt, f64, f32);
    atomicAdd(&a.flops[0], f64 * nval);
    atomicAdd(&a.flops[1], f32 * nval);
  }
}

}  // namespace ODV_VARIANT_NS
}  // namespace odv

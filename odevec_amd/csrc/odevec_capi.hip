// C ABI of libodevec (declarations and reference citations in include/odevec.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/odevec.h"
#include "odevec_registry.h"

namespace odv {

std::vector<const NetEntry*>& registry() {
  static std::vector<const NetEntry*> r;
  return r;
}

std::vector<const BundleEntry*>& bundles() {
  static std::vector<const BundleEntry*> r;
  return r;
}

__global__ void k_setq(unsigned long long* q, unsigned long long v) { *q = v; }
__global__ void k_setu(unsigned* p, unsigned v) { *p = v; }
__global__ void k_fill_int(int* p, long long n, int v) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}
__global__ void k_zero_u(unsigned* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

__global__ void k_fill_f64(double* p, double v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace odv

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(ODEVEC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// "NET/ORDERING/PACKAGING" selects the default build of a network,
// "NET/ORDERING/PACKAGING+REPRO" its reproducible build.
const odv::NetEntry* find_net(const char* id) {
  if (!id) return nullptr;
  std::string s(id);
  int repro = 0;
  const std::string suf = "+REPRO";
  if (s.size() > suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0) {
    repro = 1;
    s.resize(s.size() - suf.size());
  }
  for (const odv::NetEntry* e : odv::registry())
    if (e->repro == repro && s == e->id) return e;
  return nullptr;
}

// A batch segment's network in the bundle's own build (a reproducible
// bundle's FP64 passes and cost ranking use the reproducible kernels)
const odv::NetEntry* find_net_build(const char* id, int repro) {
  if (!id) return nullptr;
  for (const odv::NetEntry* e : odv::registry())
    if (e->repro == repro && std::strcmp(id, e->id) == 0) return e;
  return nullptr;
}

}  // namespace

struct odevec_s {
  const odv::NetEntry* net = nullptr;
  int64_t nvector = 0;
  odevec_opts_t opts{};
  double rtol = 1e-6, atol = 1e-10, dt_min = (double)1e-10f;
  int check_neg = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int64_t cap = 0;
  double* d_y = nullptr;
  double* d_t0 = nullptr;
  double* d_ts = nullptr;
  double* d_dt0 = nullptr;
  double* d_prm = nullptr;
  double* d_dtout = nullptr;
  int* d_status = nullptr;
  int* d_cnt = nullptr;
  int8_t* d_mask = nullptr;
  double* d_scal = nullptr;    // [0..2] uniform t0, t_stop, dt0; [3..] default params; [..] flops
  unsigned long long* d_queue = nullptr;   // 64 queue counters, one per in-flight persistent launch
  double* d_ws = nullptr;            // rate workspace of the per-cell kernels
  int64_t ws_cap = 0;                // doubles
  int qslot = 0;
  int nscal = 0;
  double* d_hist = nullptr;          // caller's warm-restart history (odevec_set_history), or null
  int64_t hist_cells = 0;
  long long* d_times = nullptr;      // caller's per-cell start/end ticks (odevec_set_timeline), or null
  int64_t times_cells = 0;
  int* d_order = nullptr;            // persistent schedule: queue position -> cell
  unsigned char* d_bin = nullptr;    // per-cell cost bin
  unsigned* d_count = nullptr;       // COST_BINS counters / offsets
  int64_t order_cap = 0;
  int promote_cf = 0;                // mixed form: FP64 fallback after this many convergence failures (0: off)
  double inv_tol = ODEVEC_MIXED_INVARIANT_TOL;   // mixed form: invariant test of the corrector (0: off)
  int promote_inv = ODEVEC_FALLBACK_INVARIANT_FAILS;   // after this many failed tests: FP64 fallback, or no test (0: neither)
  int mixed_rms = 0;                 // mixed form: the reference's RMS corrector test instead of the max norm
  int* d_promo = nullptr;            // cells handed to the FP64 pass
  unsigned* d_promo_n = nullptr;     // their count
  int64_t promo_cap = 0;
  double* d_ho = nullptr;            // their hand-off records (NetEntry::handoff_doubles per cell)
  int64_t ho_cap = 0;
  double* d_ws2 = nullptr;           // persistent mode with the fallback: the FP64 lanes' workspace,
  int64_t ws2_cap = 0;
  unsigned* d_done = nullptr;        // the count of finished queue wavefronts,
  hipStream_t fb_stream = nullptr;   // the stream the FP64 lanes run on beside the queue kernel,
  hipEvent_t fb_fork = nullptr, fb_join = nullptr;   // and its fork / join with the caller's stream
  // persistent mode, queue in cost order: the inputs and outputs in queue
  // order (k_gather / k_scatter), g_cap cells
  double* d_gy = nullptr;
  double* d_gts = nullptr;
  double* d_gdt = nullptr;
  double* d_gt0 = nullptr;
  double* d_gprm = nullptr;
  double* d_gdto = nullptr;
  int* d_gst = nullptr;
  int* d_gcnt = nullptr;
  int* d_ginv = nullptr;             // queue position of each cell
  int64_t g_cap = 0;
  std::vector<int> h_status;
  std::vector<int> h_cnt;
};

namespace {

void free_bufs(odevec_s* h) {
  void* ps[] = {h->d_y, h->d_t0, h->d_ts, h->d_dt0, h->d_prm, h->d_dtout, h->d_status, h->d_cnt, h->d_mask};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  h->d_y = h->d_t0 = h->d_ts = h->d_dt0 = h->d_prm = h->d_dtout = nullptr;
  h->d_status = h->d_cnt = nullptr;
  h->d_mask = nullptr;
  h->cap = 0;
}

int ensure_cap(odevec_s* h, int64_t n) {
  if (n <= h->cap) return ODEVEC_OK;
  free_bufs(h);
  const int64_t neq = h->net->neq, np = h->net->nparam;
  hipError_t e = hipSuccess;
#define ODV_ALLOC(p, bytes)                                        \
  if (e == hipSuccess) e = hipMalloc((void**)&(p), (size_t)(bytes));
  ODV_ALLOC(h->d_y, neq * n * 8);
  ODV_ALLOC(h->d_t0, n * 8);
  ODV_ALLOC(h->d_ts, n * 8);
  ODV_ALLOC(h->d_dt0, n * 8);
  ODV_ALLOC(h->d_prm, (np > 0 ? np : 1) * n * 8);
  ODV_ALLOC(h->d_dtout, n * 8);
  ODV_ALLOC(h->d_status, n * 4);
  ODV_ALLOC(h->d_cnt, 8 * n * 4);
  ODV_ALLOC(h->d_mask, n);
#undef ODV_ALLOC
  if (e != hipSuccess) {
    free_bufs(h);
    return fail(ODEVEC_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
  }
  h->cap = n;
  return ODEVEC_OK;
}

struct Req {
  int64_t ncells;
  const double* t0; int64_t t0_s;
  const double* ts; int64_t ts_s;
  const double* dt0; int64_t dt_s;
  double* y;
  const double* prm; int64_t ps_cell;
  const int8_t* mask;
  int* status;
  double* dt_out;
  int* counters;
  double* flops;
  hipStream_t stream;
};

int launch(odevec_s* h, const Req& r) {
  odv::KArgs a{};
  a.y = r.y;
  a.ncells = r.ncells;
  a.prm = r.prm;
  a.ps_cell = r.ps_cell;
  a.ps_plane = (r.ps_cell == 1) ? r.ncells : 1;
  a.t0 = r.t0; a.t0_s = r.t0_s;
  a.tstop = r.ts; a.ts_s = r.ts_s;
  a.dt0 = r.dt0; a.dt_s = r.dt_s;
  a.mask = reinterpret_cast<const signed char*>(r.mask);
  a.status = r.status;
  a.dt_out = r.dt_out;
  a.counters = r.counters;
  a.flops = r.flops;
  if (h->d_times) {
    if (r.ncells != h->times_cells)
      return fail(ODEVEC_EINVAL, "the solve covers " + std::to_string(r.ncells) + " cells; the timeline holds " +
                                     std::to_string(h->times_cells));
    a.times = h->d_times;
  }
  if (h->d_hist) {
    if (r.ncells != h->hist_cells)
      return fail(ODEVEC_EINVAL, "the solve covers " + std::to_string(r.ncells) + " cells; the history holds " +
                                     std::to_string(h->hist_cells));
    a.hist = h->d_hist;
  }
  h->qslot = (h->qslot + 1) % 64;
  a.queue = h->d_queue + h->qslot;
  a.ctl.rtol = h->rtol;
  a.ctl.atol = h->atol;
  a.ctl.dt_min = h->dt_min;
  a.ctl.check_negatives = h->check_neg;
  a.ctl.max_attempts = h->opts.max_attempts > 0 ? h->opts.max_attempts : 1000000;
  a.ctl.inv_tol = h->inv_tol;
  a.ctl.promote_inv = h->promote_inv;
  a.ctl.mixed_rms = h->mixed_rms;
  int grp = h->opts.group;
  if (grp <= 0) grp = (int)(h->nvector < 256 ? h->nvector : 256);
  if (grp < 1) grp = 1;
  a.group = grp;
  odv::Plan p{};
  hipError_t e = h->net->plan(r.ncells, h->opts.mode, h->opts.precision, h->opts.block, h->opts.grid, grp, &p);
  if (e != hipSuccess) return hip_fail(e, "launch plan");
  // the FP64 fallback: cells that reach promote_cf convergence failures stop
  // (their input state untouched, their integrator state in a hand-off
  // record) and the FP64 persistent kernel continues them over the list of
  // such cells (second launch); on an FP64 handle the cells only move
  // between the two launches, with unchanged results
  const bool promote = h->promote_cf > 0 && (h->opts.mode == ODEVEC_PERCELL || h->opts.mode == ODEVEC_PERSISTENT);
  // the fallback's list and the cost-ordered queue hold cell indices as int
  if ((promote || (h->opts.mode == ODEVEC_PERSISTENT && h->opts.schedule == 1)) && r.ncells > 0x7fffffffLL)
    return fail(ODEVEC_EINVAL, "the FP64 fallback and the cost-ordered queue take at most 2^31 - 1 cells per solve");
  odv::Plan p2{};
  // persistent mode: the FP64 lanes run beside the queue kernel, on a stream
  // of the handle, in 1 in FB_SHARE of the machine's wavefront slots; the pass
  // after both only takes what they left
  const bool beside = promote && h->opts.mode == ODEVEC_PERSISTENT;
  long long nfb = 0;
  int nprod = 0;
  if (promote) {
    e = h->net->plan(r.ncells, ODEVEC_PERSISTENT, ODEVEC_FP64, h->opts.block, h->opts.grid, grp, &p2);
    if (e != hipSuccess) return hip_fail(e, "launch plan (FP64 pass)");
    if (beside) {
      const long long wpb = p.block / 64, waves = p.grid * wpb;
      nfb = waves / odv::FB_SHARE > 0 ? waves / odv::FB_SHARE : 1;
      const long long keep = (waves - nfb) / wpb;
      p.grid = keep > 0 ? keep : 1;
      nprod = (int)(p.grid * wpb);
      long long ws64 = nfb * h->net->ws64_per_wave;
      if (p2.ws_doubles > ws64) ws64 = p2.ws_doubles;
      if (ws64 > h->ws2_cap) {
        if (h->d_ws2) (void)hipFree(h->d_ws2);
        h->d_ws2 = nullptr;
        h->ws2_cap = 0;
        if ((e = hipMalloc((void**)&h->d_ws2, (size_t)ws64 * 8)) != hipSuccess)
          return fail(ODEVEC_ENOMEM, std::string("fallback workspace allocation failed: ") + hipGetErrorString(e));
        h->ws2_cap = ws64;
      }
      if (!h->d_done && (e = hipMalloc((void**)&h->d_done, sizeof(unsigned))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback counter allocation failed: ") + hipGetErrorString(e));
      if (!h->fb_stream && (e = hipStreamCreateWithFlags(&h->fb_stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "fallback stream");
      if (!h->fb_fork && (e = hipEventCreateWithFlags(&h->fb_fork, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "fallback event");
      if (!h->fb_join && (e = hipEventCreateWithFlags(&h->fb_join, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "fallback event");
    } else if (p2.ws_doubles > p.ws_doubles) {
      p.ws_doubles = p2.ws_doubles;
    }
    if (r.ncells > h->promo_cap) {
      if (h->d_promo) (void)hipFree(h->d_promo);
      h->d_promo = nullptr;
      h->promo_cap = 0;
      if ((e = hipMalloc((void**)&h->d_promo, (size_t)r.ncells * sizeof(int))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback list allocation failed: ") + hipGetErrorString(e));
      h->promo_cap = r.ncells;
    }
    if (!h->d_promo_n && (e = hipMalloc((void**)&h->d_promo_n, sizeof(unsigned))) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("fallback list allocation failed: ") + hipGetErrorString(e));
    const int64_t need = r.ncells * (int64_t)h->net->handoff_doubles;
    if (need > h->ho_cap) {
      if (h->d_ho) (void)hipFree(h->d_ho);
      h->d_ho = nullptr;
      h->ho_cap = 0;
      if ((e = hipMalloc((void**)&h->d_ho, (size_t)need * sizeof(double))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback record allocation failed: ") + hipGetErrorString(e));
      h->ho_cap = need;
    }
  }
  if (p.ws_doubles > h->ws_cap) {
    // hipFree waits for the device, so no earlier launch still reads the old workspace
    if (h->d_ws) (void)hipFree(h->d_ws);
    h->d_ws = nullptr;
    h->ws_cap = 0;
    e = hipMalloc((void**)&h->d_ws, (size_t)p.ws_doubles * 8);
    if (e != hipSuccess) {
      h->d_ws = nullptr;
      return fail(ODEVEC_ENOMEM, std::string("rate workspace allocation failed: ") + hipGetErrorString(e));
    }
    h->ws_cap = p.ws_doubles;
  }
  a.ws = h->d_ws;
  // cost-ordered queue when the cells outnumber the launched lanes
  if (h->opts.mode == ODEVEC_PERSISTENT && h->opts.schedule == 1 && r.ncells > p.grid * p.block) {
    if (r.ncells > h->order_cap) {
      if (h->d_order) (void)hipFree(h->d_order);
      if (h->d_bin) (void)hipFree(h->d_bin);
      h->d_order = nullptr;
      h->d_bin = nullptr;
      h->order_cap = 0;
      if ((e = hipMalloc((void**)&h->d_order, (size_t)r.ncells * sizeof(int))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_bin, (size_t)r.ncells)) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("schedule allocation failed: ") + hipGetErrorString(e));
      h->order_cap = r.ncells;
    }
    if (!h->d_count && (e = hipMalloc((void**)&h->d_count, odv::COST_BINS * sizeof(unsigned))) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("schedule allocation failed: ") + hipGetErrorString(e));
    if ((e = h->net->order(a, h->d_bin, h->d_count, h->d_order, r.stream)) != hipSuccess)
      return hip_fail(e, "cost order");
    a.order = h->d_order;
  }
  // the solve on copies in queue order (coalesced state accesses in the
  // persistent loop), unless per-cell side arrays indexed by cell are attached
  odv::KArgs src = a;
  const bool gather = a.order && !a.hist && !a.times && !a.mask;
  if (gather) {
    const int64_t n = r.ncells, neq = h->net->neq, np = h->net->nparam;
    if (n > h->g_cap) {
      void* ps[] = {h->d_gy, h->d_gts, h->d_gdt, h->d_gt0, h->d_gprm, h->d_gdto, h->d_gst, h->d_gcnt, h->d_ginv};
      for (void* q : ps)
        if (q) (void)hipFree(q);
      h->d_gy = h->d_gts = h->d_gdt = h->d_gt0 = h->d_gprm = h->d_gdto = nullptr;
      h->d_gst = h->d_gcnt = h->d_ginv = nullptr;
      h->g_cap = 0;
      if ((e = hipMalloc((void**)&h->d_gy, (size_t)(neq * n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gts, (size_t)(n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gdt, (size_t)(n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gt0, (size_t)(n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gprm, (size_t)((np > 0 ? np : 1) * n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gdto, (size_t)(n * 8))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gst, (size_t)(n * 4))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_gcnt, (size_t)(8 * n * 4))) != hipSuccess ||
          (e = hipMalloc((void**)&h->d_ginv, (size_t)(n * 4))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("queue-order copies allocation failed: ") + hipGetErrorString(e));
      h->g_cap = n;
    }
    double* gts = a.ts_s ? h->d_gts : nullptr;
    double* gdt = a.dt_s ? h->d_gdt : nullptr;
    double* gt0 = a.t0_s ? h->d_gt0 : nullptr;
    double* gprm = (np > 0 && a.ps_cell) ? h->d_gprm : nullptr;
    hipLaunchKernelGGL(odv::k_invert<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, r.stream, (long long)n,
                       (const int*)h->d_order, h->d_ginv);
    hipLaunchKernelGGL(odv::k_gather<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, r.stream, a, (int)neq,
                       (int)np, (const int*)h->d_ginv, h->d_gy, gts, gdt, gt0, gprm);
    a.y = h->d_gy;
    if (gts) { a.tstop = gts; a.ts_s = 1; }
    if (gdt) { a.dt0 = gdt; a.dt_s = 1; }
    if (gt0) { a.t0 = gt0; a.t0_s = 1; }
    if (gprm) { a.prm = gprm; a.ps_plane = n; a.ps_cell = 1; }
    if (a.status) a.status = h->d_gst;
    if (a.dt_out) a.dt_out = h->d_gdto;
    if (a.counters) a.counters = h->d_gcnt;
    a.order = nullptr;
    a.spread = 1;
  }
  auto scatter = [&]() {
    if (!gather) return;
    hipLaunchKernelGGL(odv::k_scatter<0>, dim3((unsigned)((r.ncells + 255) / 256)), dim3(256), 0, r.stream,
                       (long long)r.ncells, (int)h->net->neq, (const int*)h->d_order, (const double*)h->d_gy,
                       (const int*)h->d_gst, (const double*)h->d_gdto, (const int*)h->d_gcnt, src.y, src.status,
                       src.dt_out, src.counters);
  };
  if (beside) {
    a.ctl.promote_cf = h->promote_cf;
    a.promo = h->d_promo;
    a.promo_n = h->d_promo_n;
    a.ho = h->d_ho;
    a.ho_mode = 1;
    a.fb_done = h->d_done;
    // the FP64 lanes: the list as their queue, claimed through their own counter
    odv::KArgs b = a;
    b.ctl.promote_cf = 0;
    b.promo = nullptr;
    b.promo_n = nullptr;
    b.fb_done = nullptr;
    b.order = h->d_promo;
    b.qlen = h->d_promo_n;
    b.ws = h->d_ws2;
    b.ho_mode = 2;
    h->qslot = (h->qslot + 1) % 64;
    b.queue = h->d_queue + h->qslot;
    // the list's entries read -1 until written; the counters start at 0
    hipLaunchKernelGGL(odv::k_fill_int, dim3((unsigned)((r.ncells + 255) / 256)), dim3(256), 0, r.stream, h->d_promo,
                       (long long)r.ncells, -1);
    hipLaunchKernelGGL(odv::k_setu, dim3(1), dim3(1), 0, r.stream, h->d_promo_n, 0u);
    hipLaunchKernelGGL(odv::k_setu, dim3(1), dim3(1), 0, r.stream, h->d_done, 0u);
    hipLaunchKernelGGL(odv::k_setq, dim3(1), dim3(1), 0, r.stream, b.queue, 0ull);
    if ((e = hipEventRecord(h->fb_fork, r.stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(h->fb_stream, h->fb_fork, 0)) != hipSuccess)
      return hip_fail(e, "fallback fork");
    e = h->net->launch(a, ODEVEC_PERSISTENT, h->opts.precision, p, r.stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    e = h->net->launch_fbc(b, h->d_done, nprod, nfb, h->fb_stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch (FP64 lanes)");
    if ((e = hipEventRecord(h->fb_join, h->fb_stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(r.stream, h->fb_join, 0)) != hipSuccess)
      return hip_fail(e, "fallback join");
    // what the FP64 lanes did not take (normally nothing): the FP64 pass
    odv::KArgs a2 = b;
    h->qslot = (h->qslot + 1) % 64;
    a2.queue = h->d_queue + h->qslot;
    e = h->net->launch(a2, ODEVEC_PERSISTENT, ODEVEC_FP64, p2, r.stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch (FP64 pass)");
    scatter();
    return ODEVEC_OK;
  }
  if (promote) {
    a.ctl.promote_cf = h->promote_cf;
    a.promo = h->d_promo;
    a.promo_n = h->d_promo_n;
    a.ho = h->d_ho;
    a.ho_mode = 1;
    hipLaunchKernelGGL(odv::k_setu, dim3(1), dim3(1), 0, r.stream, h->d_promo_n, 0u);
  }
  e = h->net->launch(a, h->opts.mode, h->opts.precision, p, r.stream);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  if (promote) {
    odv::KArgs a2 = a;
    a2.ctl.promote_cf = 0;
    a2.promo = nullptr;
    a2.promo_n = nullptr;
    a2.order = h->d_promo;
    a2.qlen = h->d_promo_n;
    a2.ho_mode = 2;
    h->qslot = (h->qslot + 1) % 64;
    a2.queue = h->d_queue + h->qslot;
    e = h->net->launch(a2, ODEVEC_PERSISTENT, ODEVEC_FP64, p2, r.stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch (FP64 pass)");
  }
  scatter();
  return ODEVEC_OK;
}

int check_status(odevec_s* h, const int* hs, int64_t n, odevec_stats_t* st) {
  int64_t ok = 0, bad = 0, skip = 0;
  int first_bad = -1;
  int64_t first_i = -1;
  for (int64_t i = 0; i < n; ++i) {
    if (hs[i] == ODEVEC_STATUS_OK) ok++;
    else if (hs[i] == ODEVEC_STATUS_SKIPPED) skip++;
    else {
      if (first_bad < 0) { first_bad = hs[i]; first_i = i; }
      bad++;
    }
  }
  if (st) { st->cells_ok = ok; st->cells_failed = bad; st->cells_skipped = skip; }
  if (bad) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%lld of %lld cells failed (first: cell %lld, status %d)",
                  (long long)bad, (long long)n, (long long)first_i, first_bad);
    if (h->opts.abort_on_error) {
      std::fprintf(stderr, "ODEVEC: %s. Aborting.\n", buf);
      std::abort();
    }
    return fail(ODEVEC_ESOLVE, buf);
  }
  return ODEVEC_OK;
}

}  // namespace

extern "C" {

int32_t odevec_abi_version(void) { return ODEVEC_ABI_VERSION; }

const char* odevec_last_error(void) { return g_err.c_str(); }

int32_t odevec_network_count(void) { return (int32_t)odv::registry().size(); }

const char* odevec_network_name(int32_t i) {
  static std::vector<std::string> names;
  if (i < 0 || i >= (int32_t)odv::registry().size()) return nullptr;
  if (names.size() != odv::registry().size()) {
    names.clear();
    for (const odv::NetEntry* e : odv::registry()) names.push_back(std::string(e->id) + (e->repro ? "+REPRO" : ""));
  }
  return names[(size_t)i].c_str();
}

int32_t odevec_network_info(const char* network, odevec_network_info_t* info) {
  const odv::NetEntry* e = find_net(network);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (!info) return fail(ODEVEC_EINVAL, "info is NULL");
  info->neq = e->neq; info->nnz = e->nnz; info->nrate = e->nrate; info->nparam = e->nparam;
  info->rate_kind = e->rate_kind; info->tgas_index = e->tgas;
  info->flops_rhs = e->flops_rhs; info->flops_jac = e->flops_jac; info->flops_lu = e->flops_lu;
  info->flops_solve = e->flops_solve; info->flops_lu_r = e->flops_lu_r; info->flops_solve_r = e->flops_solve_r;
  return ODEVEC_OK;
}

int32_t odevec_history_planes(const char* network, int64_t* planes) {
  const odv::NetEntry* e = find_net(network);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (!planes) return fail(ODEVEC_EINVAL, "planes is NULL");
  if (!e->lane_ok && !e->coop_ok)
    return fail(ODEVEC_EINVAL, std::string(e->id) + ": warm restart needs the per-lane or the cooperative kernels");
  *planes = e->hist_planes;
  return ODEVEC_OK;
}

int32_t odevec_set_history(odevec_t* h, double* d_hist, int64_t ncells) {
  if (!h) return fail(ODEVEC_EINVAL, "NULL handle");
  if (!d_hist) {
    h->d_hist = nullptr;
    h->hist_cells = 0;
    return ODEVEC_OK;
  }
  if (ncells <= 0) return fail(ODEVEC_EINVAL, "ncells must be > 0");
  const bool coop_only = h->opts.mode == ODEVEC_COOP && !h->net->lane_ok;
  if (h->opts.mode != ODEVEC_PERCELL && h->opts.mode != ODEVEC_PERSISTENT && !coop_only)
    return fail(ODEVEC_EINVAL, "warm restart needs the per-cell or persistent mode (one step control per cell), "
                               "or the cooperative mode for a network too large for one lane");
  h->d_hist = d_hist;
  h->hist_cells = ncells;
  return ODEVEC_OK;
}

int32_t odevec_set_timeline(odevec_t* h, int64_t* d_times, int64_t ncells, int64_t* ticks_per_second) {
  if (!h) return fail(ODEVEC_EINVAL, "NULL handle");
  if (ticks_per_second) {
    int khz = 0;
    hipError_t e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device);
    if (e != hipSuccess) return hip_fail(e, "wall clock rate");
    *ticks_per_second = (int64_t)khz * 1000;
  }
  if (!d_times) {
    h->d_times = nullptr;
    h->times_cells = 0;
    return ODEVEC_OK;
  }
  if (ncells <= 0) return fail(ODEVEC_EINVAL, "ncells must be > 0");
  if (h->opts.mode == ODEVEC_LOCKSTEP)
    return fail(ODEVEC_EINVAL, "the timeline records per-cell, persistent and cooperative solves");
  h->d_times = reinterpret_cast<long long*>(d_times);
  h->times_cells = ncells;
  return ODEVEC_OK;
}

int32_t odevec_kernel_info(const char* network, int32_t mode, int32_t precision, int32_t block,
                           odevec_kernel_info_t* info) {
  const odv::NetEntry* e = find_net(network);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (!info) return fail(ODEVEC_EINVAL, "info is NULL");
  odv::KernelInfo k{};
  hipError_t he = e->info(mode, precision, block, &k);
  if (he != hipSuccess) return hip_fail(he, "kernel info");
  info->vgprs = k.regs; info->spill_bytes = k.spill_bytes;
  info->blocks_per_cu = k.blocks_per_cu; info->block = k.block;
  return ODEVEC_OK;
}

void odevec_default_opts(odevec_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->precision = ODEVEC_FP64;
  o->mode = ODEVEC_PERCELL;
}

int32_t odevec_create(const char* network, int64_t nvector, const odevec_opts_t* opts, odevec_t** out) {
  if (!out) return fail(ODEVEC_EINVAL, "out is NULL");
  *out = nullptr;
  const odv::NetEntry* e = find_net(network);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (nvector < 1) return fail(ODEVEC_EINVAL, "nvector must be >= 1");
  odevec_opts_t o;
  odevec_default_opts(&o);
  if (opts) o = *opts;
  if (o.precision != ODEVEC_FP64 && o.precision != ODEVEC_MIXED) return fail(ODEVEC_EINVAL, "bad precision");
  if (o.mode < ODEVEC_PERCELL || o.mode > ODEVEC_COOP) return fail(ODEVEC_EINVAL, "bad mode");
  if (o.schedule < 0 || o.schedule > 1) return fail(ODEVEC_EINVAL, "schedule must be 0 (index order) or 1 (costliest first)");
  if (o.mode == ODEVEC_COOP && !e->coop_ok)
    return fail(ODEVEC_EINVAL, "the cooperative mode needs a CSC network of at most 64 integrated species");
  if (o.mode != ODEVEC_COOP && !e->lane_ok)
    return fail(ODEVEC_EINVAL, "more than 32 integrated species: only the cooperative mode (ODEVEC_COOP)");
  if (o.group < 0 || o.group > 256) return fail(ODEVEC_EINVAL, "group must be in 0..256");
  if (o.block != 0 && (o.block < 64 || o.block > 256 || o.block % 64)) return fail(ODEVEC_EINVAL, "block must be 64..256, a multiple of 64");
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev < 1) return fail(ODEVEC_EDEVICE, "no HIP device available");
  odevec_s* h = new odevec_s();
  h->net = e;
  h->nvector = nvector;
  h->opts = o;
  hipGetDevice(&h->device);
  if ((he = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess ||
      (he = hipEventCreate(&h->e0)) != hipSuccess || (he = hipEventCreate(&h->e1)) != hipSuccess) {
    odevec_destroy(h);
    return hip_fail(he, "stream/event creation");
  }
  const int np = e->nparam > 0 ? e->nparam : 1;
  h->nscal = 3 + np + 2;
  if ((he = hipMalloc((void**)&h->d_scal, h->nscal * 8)) != hipSuccess ||
      (he = hipMalloc((void**)&h->d_queue, 64 * sizeof(unsigned long long))) != hipSuccess) {
    odevec_destroy(h);
    return fail(ODEVEC_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(he));
  }
  std::vector<double> sc(h->nscal, 0.0);
  sc[2] = 1e-10;
  e->param_defaults(sc.data() + 3);
  if ((he = hipMemcpy(h->d_scal, sc.data(), h->nscal * 8, hipMemcpyHostToDevice)) != hipSuccess) {
    odevec_destroy(h);
    return hip_fail(he, "upload defaults");
  }
  int rc = ensure_cap(h, nvector);
  if (rc != ODEVEC_OK) {
    std::string m = g_err;
    odevec_destroy(h);
    return fail(rc, m);
  }
  *out = h;
  return ODEVEC_OK;
}

int32_t odevec_set_tolerances(odevec_t* h, double rtol, double atol) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  if (!(rtol > 0.0) || !(atol >= 0.0) || (atol == 0.0 && rtol == 0.0))
    return fail(ODEVEC_EINVAL, "need rtol > 0 and atol >= 0");
  h->rtol = rtol;
  h->atol = atol;
  return ODEVEC_OK;
}

int32_t odevec_set_dt_min(odevec_t* h, double dt_min) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  h->dt_min = dt_min;
  return ODEVEC_OK;
}

int32_t odevec_set_check_negatives(odevec_t* h, int32_t flag) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  h->check_neg = flag ? 1 : 0;
  return ODEVEC_OK;
}

int32_t odevec_set_fp64_fallback(odevec_t* h, int32_t convfail_limit) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  if (convfail_limit < 0) return fail(ODEVEC_EINVAL, "convfail_limit must be >= 0");
  h->promote_cf = convfail_limit;
  return ODEVEC_OK;
}

int32_t odevec_set_mixed_convergence_test(odevec_t* h, int32_t test) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  if (test != ODEVEC_MIXED_TEST_MAX && test != ODEVEC_MIXED_TEST_RMS)
    return fail(ODEVEC_EINVAL, "test must be ODEVEC_MIXED_TEST_MAX (0) or ODEVEC_MIXED_TEST_RMS (1)");
  h->mixed_rms = test == ODEVEC_MIXED_TEST_RMS ? 1 : 0;
  return ODEVEC_OK;
}

int32_t odevec_set_invariant_test(odevec_t* h, double tolerances, int32_t fallback_fails) {
  if (!h) return fail(ODEVEC_EINVAL, "handle is NULL");
  if (!(tolerances >= 0.0)) return fail(ODEVEC_EINVAL, "tolerances must be >= 0");
  if (fallback_fails < 0) return fail(ODEVEC_EINVAL, "fallback_fails must be >= 0");
  h->inv_tol = tolerances;
  h->promote_inv = fallback_fails;
  return ODEVEC_OK;
}

int32_t odevec_solve(odevec_t* h, double* time, double* dt, double t_stop, double* y, const int8_t* mask,
                     int32_t* status) {
  if (!h || !time || !dt || !y) return fail(ODEVEC_EINVAL, "NULL argument");
  const int64_t n = h->nvector;
  const int64_t neq = h->net->neq;
  hipStream_t st = h->stream;
  hipError_t e;
  double sc[3] = {*time, t_stop, *dt};
  if ((e = hipMemcpyAsync(h->d_scal, sc, 24, hipMemcpyHostToDevice, st)) != hipSuccess) return hip_fail(e, "H2D");
  double* dy = h->opts.device_ptrs ? y : h->d_y;
  if (!h->opts.device_ptrs &&
      (e = hipMemcpyAsync(h->d_y, y, (size_t)(neq * n * 8), hipMemcpyHostToDevice, st)) != hipSuccess)
    return hip_fail(e, "H2D y");
  const int8_t* dmask = nullptr;
  if (mask) {
    if (h->opts.device_ptrs) dmask = mask;
    else {
      if ((e = hipMemcpyAsync(h->d_mask, mask, (size_t)n, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "H2D mask");
      dmask = h->d_mask;
    }
  }
  Req r{};
  r.ncells = n;
  r.t0 = h->d_scal; r.t0_s = 0;
  r.ts = h->d_scal + 1; r.ts_s = 0;
  r.dt0 = h->d_scal + 2; r.dt_s = 0;
  r.y = dy;
  r.prm = h->d_scal + 3; r.ps_cell = 0;
  r.mask = dmask;
  r.status = h->d_status;
  r.dt_out = h->d_dtout;
  r.stream = st;
  hipEventRecord(h->e0, st);
  int rc = launch(h, r);
  hipEventRecord(h->e1, st);
  if (rc != ODEVEC_OK) return rc;
  if (!h->opts.device_ptrs &&
      (e = hipMemcpyAsync(y, h->d_y, (size_t)(neq * n * 8), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "D2H y");
  h->h_status.resize((size_t)n);
  double dt0 = 0.0;
  if ((e = hipMemcpyAsync(h->h_status.data(), h->d_status, (size_t)n * 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipMemcpyAsync(&dt0, h->d_dtout, 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "D2H status");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "solve");
  if (status) std::memcpy(status, h->h_status.data(), (size_t)n * 4);
  *time = t_stop;
  *dt = dt0;
  return check_status(h, h->h_status.data(), n, nullptr);
}

int32_t odevec_solve_cells(odevec_t* h, int64_t ncells, const double* t0, const double* t_stop,
                           const double* dt0, double* y, const double* params, int32_t* status,
                           double* dt_out, odevec_stats_t* stats) {
  if (!h || !t_stop || !y) return fail(ODEVEC_EINVAL, "NULL argument");
  if (ncells < 0) return fail(ODEVEC_EINVAL, "ncells < 0");
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (ncells == 0) return ODEVEC_OK;
  const bool dev = h->opts.device_ptrs != 0;
  int rc = ensure_cap(h, ncells);
  if (rc != ODEVEC_OK) return rc;
  const int64_t neq = h->net->neq, np = h->net->nparam;
  hipStream_t st = h->stream;
  hipError_t e;
  auto h2d = [&](double* dst, const double* src, int64_t cnt) {
    return hipMemcpyAsync(dst, src, (size_t)cnt * 8, hipMemcpyHostToDevice, st);
  };
  Req r{};
  r.ncells = ncells;
  r.stream = st;
  if (dev) {
    r.y = y;
    r.ts = t_stop; r.ts_s = 1;
  } else {
    if ((e = h2d(h->d_y, y, neq * ncells)) != hipSuccess) return hip_fail(e, "H2D y");
    if ((e = h2d(h->d_ts, t_stop, ncells)) != hipSuccess) return hip_fail(e, "H2D t_stop");
    r.y = h->d_y;
    r.ts = h->d_ts; r.ts_s = 1;
  }
  if (t0) {
    if (dev) { r.t0 = t0; }
    else { if ((e = h2d(h->d_t0, t0, ncells)) != hipSuccess) return hip_fail(e, "H2D t0"); r.t0 = h->d_t0; }
    r.t0_s = 1;
  } else { r.t0 = h->d_scal; r.t0_s = 0; }
  double zero_and_hint[3] = {0.0, 0.0, 1e-10};
  if ((e = h2d(h->d_scal, zero_and_hint, 3)) != hipSuccess) return hip_fail(e, "H2D scalars");
  if (dt0) {
    if (dev) { r.dt0 = dt0; }
    else { if ((e = h2d(h->d_dt0, dt0, ncells)) != hipSuccess) return hip_fail(e, "H2D dt0"); r.dt0 = h->d_dt0; }
    r.dt_s = 1;
  } else { r.dt0 = h->d_scal + 2; r.dt_s = 0; }
  if (params && np > 0) {
    if (dev) { r.prm = params; }
    else { if ((e = h2d(h->d_prm, params, np * ncells)) != hipSuccess) return hip_fail(e, "H2D params"); r.prm = h->d_prm; }
    r.ps_cell = 1;
  } else { r.prm = h->d_scal + 3; r.ps_cell = 0; }
  r.status = h->d_status;
  r.dt_out = (dt_out && dev) ? dt_out : h->d_dtout;
  r.counters = stats ? h->d_cnt : nullptr;
  double* dfl = h->d_scal + 3 + (np > 0 ? np : 1);
  if (stats) {
    hipLaunchKernelGGL(odv::k_fill_f64, dim3(1), dim3(64), 0, st, dfl, 0.0, 2LL);
    r.flops = dfl;
  }
  hipEventRecord(h->e0, st);
  rc = launch(h, r);
  hipEventRecord(h->e1, st);
  if (rc != ODEVEC_OK) return rc;
  if (!dev && (e = hipMemcpyAsync(y, h->d_y, (size_t)(neq * ncells * 8), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "D2H y");
  if (dt_out && !dev &&
      (e = hipMemcpyAsync(dt_out, h->d_dtout, (size_t)ncells * 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "D2H dt");
  h->h_status.resize((size_t)ncells);
  if ((e = hipMemcpyAsync(h->h_status.data(), h->d_status, (size_t)ncells * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return hip_fail(e, "D2H status");
  double fl[2] = {0.0, 0.0};
  if (stats) {
    h->h_cnt.resize((size_t)(8 * ncells));
    if ((e = hipMemcpyAsync(h->h_cnt.data(), h->d_cnt, (size_t)(8 * ncells * 4), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipMemcpyAsync(fl, dfl, 16, hipMemcpyDeviceToHost, st)) != hipSuccess)
      return hip_fail(e, "D2H stats");
  }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "solve");
  if (status) std::memcpy(status, h->h_status.data(), (size_t)ncells * 4);
  if (stats) {
    int64_t* dst = &stats->attempts;
    for (int m = 0; m < 8; ++m) {
      int64_t s = 0;
      const int* c = h->h_cnt.data() + (size_t)m * ncells;
      for (int64_t i = 0; i < ncells; ++i) s += c[i];
      dst[m] = s;
    }
    stats->flops64 = fl[0];
    stats->flops32 = fl[1];
    float ms = 0.f;
    hipEventElapsedTime(&ms, h->e0, h->e1);
    stats->kernel_ms = ms;
  }
  return check_status(h, h->h_status.data(), ncells, stats);
}

namespace {

// GetRHS (what 0) / GetJac (what 1) of ncells cells through the handle's buffers
int netfun(odevec_t* h, int what, int64_t ncells, double hb, const double* y, const double* params, double* out) {
  if (!h || !y || !out) return fail(ODEVEC_EINVAL, "NULL argument");
  if (ncells < 0) return fail(ODEVEC_EINVAL, "ncells < 0");
  if (!h->net->lane_ok || h->net->jac_entries <= 0)
    return fail(ODEVEC_EINVAL, std::string(h->net->id) + ": the network functions need the per-lane code");
  if (ncells == 0) return ODEVEC_OK;
  const bool dev = h->opts.device_ptrs != 0;
  int rc = ensure_cap(h, ncells);
  if (rc != ODEVEC_OK) return rc;
  const int64_t neq = h->net->neq, np = h->net->nparam;
  const int64_t per = what == 0 ? neq : h->net->jac_entries;
  hipStream_t st = h->stream;
  hipError_t e;
  odv::KArgs a{};
  a.ncells = ncells;
  if (dev) a.y = const_cast<double*>(y);
  else {
    if ((e = hipMemcpyAsync(h->d_y, y, (size_t)(neq * ncells) * 8, hipMemcpyHostToDevice, st)) != hipSuccess)
      return hip_fail(e, "H2D y");
    a.y = h->d_y;
  }
  if (params && np > 0) {
    if (dev) a.prm = params;
    else {
      if ((e = hipMemcpyAsync(h->d_prm, params, (size_t)(np * ncells) * 8, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "H2D params");
      a.prm = h->d_prm;
    }
    a.ps_cell = 1;
    a.ps_plane = ncells;
  } else {
    a.prm = h->d_scal + 3;
    a.ps_cell = 0;
    a.ps_plane = 1;
  }
  double* d_out = out;
  if (!dev && (e = hipMalloc((void**)&d_out, (size_t)(per * ncells) * 8)) != hipSuccess)
    return fail(ODEVEC_ENOMEM, std::string("output allocation failed: ") + hipGetErrorString(e));
  e = h->net->netfun(what, a, hb, d_out, st);
  if (e == hipSuccess && !dev) e = hipMemcpyAsync(out, d_out, (size_t)(per * ncells) * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (!dev) (void)hipFree(d_out);
  return e == hipSuccess ? ODEVEC_OK : hip_fail(e, what == 0 ? "GetRHS" : "GetJac");
}

}  // namespace

int32_t odevec_get_rhs(odevec_t* h, int64_t ncells, const double* y, const double* params, double* rhs) {
  return netfun(h, 0, ncells, 0.0, y, params, rhs);
}

int32_t odevec_get_jac(odevec_t* h, int64_t ncells, double beta, double dt, const double* y, const double* params,
                       double* jac) {
  return netfun(h, 1, ncells, beta * dt, y, params, jac);
}

int32_t odevec_jac_entries(const char* network, int64_t* entries) {
  const odv::NetEntry* e = find_net(network);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (!entries) return fail(ODEVEC_EINVAL, "entries is NULL");
  if (e->jac_entries <= 0) return fail(ODEVEC_EINVAL, std::string(e->id) + ": no per-lane network code");
  *entries = e->jac_entries;
  return ODEVEC_OK;
}

int32_t odevec_solve_cells_async(odevec_t* h, int64_t ncells, const double* d_t0, int64_t t0_stride,
                                 const double* d_t_stop, int64_t t_stop_stride, const double* d_dt0,
                                 int64_t dt0_stride, double* d_y, const double* d_params,
                                 int64_t params_cell_stride, int32_t* d_status, double* d_dt_out,
                                 int32_t* d_counters, double* d_flops, void* stream) {
  if (!h || !d_t_stop || !d_y) return fail(ODEVEC_EINVAL, "NULL argument");
  if (ncells < 0) return fail(ODEVEC_EINVAL, "ncells < 0");
  if (ncells == 0) return ODEVEC_OK;
  if (t0_stride < 0 || t0_stride > 1 || t_stop_stride < 0 || t_stop_stride > 1 || dt0_stride < 0 ||
      dt0_stride > 1 || params_cell_stride < 0 || params_cell_stride > 1)
    return fail(ODEVEC_EINVAL, "strides must be 0 or 1");
  Req r{};
  r.ncells = ncells;
  r.stream = stream ? (hipStream_t)stream : h->stream;
  r.y = d_y;
  r.ts = d_t_stop; r.ts_s = t_stop_stride;
  if (d_t0) { r.t0 = d_t0; r.t0_s = t0_stride; }
  else { r.t0 = h->d_scal; r.t0_s = 0; }
  if (d_dt0) { r.dt0 = d_dt0; r.dt_s = dt0_stride; }
  else { r.dt0 = h->d_scal + 2; r.dt_s = 0; }
  if (d_params && h->net->nparam > 0) { r.prm = d_params; r.ps_cell = params_cell_stride; }
  else { r.prm = h->d_scal + 3; r.ps_cell = 0; }
  r.status = d_status;
  r.dt_out = d_dt_out;
  r.counters = d_counters;
  r.flops = d_flops;
  return launch(h, r);
}

int32_t odevec_destroy(odevec_t* h) {
  if (!h) return ODEVEC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_bufs(h);
  if (h->d_scal) (void)hipFree(h->d_scal);
  if (h->d_queue) (void)hipFree(h->d_queue);
  if (h->d_ws) (void)hipFree(h->d_ws);
  if (h->d_order) (void)hipFree(h->d_order);
  if (h->d_promo) (void)hipFree(h->d_promo);
  if (h->d_promo_n) (void)hipFree(h->d_promo_n);
  if (h->d_ho) (void)hipFree(h->d_ho);
  if (h->d_ws2) (void)hipFree(h->d_ws2);
  if (h->d_done) (void)hipFree(h->d_done);
  if (h->fb_stream) (void)hipStreamDestroy(h->fb_stream);
  if (h->fb_fork) (void)hipEventDestroy(h->fb_fork);
  if (h->fb_join) (void)hipEventDestroy(h->fb_join);
  if (h->d_bin) (void)hipFree(h->d_bin);
  if (h->d_count) (void)hipFree(h->d_count);
  {
    void* ps[] = {h->d_gy, h->d_gts, h->d_gdt, h->d_gt0, h->d_gprm, h->d_gdto, h->d_gst, h->d_gcnt, h->d_ginv};
    for (void* q : ps)
      if (q) (void)hipFree(q);
  }
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return ODEVEC_OK;
}

}  // extern "C"

/* ------------------------------------------------ solve without allocation */

// Device memory of odevec_solve_cells_noalloc beyond the caller's arrays, in
// one file-scope array per GPU: [0, 256) the persistent queue's counter,
// [256, 1024) default values (t0 = 0, the dt hint 1e-10, the network's
// parameter defaults), [1024, end) the per-wavefront workspace. In host code
// `scratch` names a host-side stand-in; its device address comes from
// hipGetSymbolAddress (per device) and only that address is used.
static __device__ __align__(256) unsigned char scratch[ODEVEC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t NOALLOC_HEAD = 1024;
constexpr int NOALLOC_MAXP = (int)((NOALLOC_HEAD - 256) / 8) - 2;
constexpr int NOALLOC_DEVICES = 64;

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_DEVICES] = {};

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= NOALLOC_DEVICES) return fail(ODEVEC_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ODEVEC_OK;
}

}  // namespace

extern "C" {

int32_t odevec_solve_cells_noalloc(const char* network, int32_t precision, int64_t ncells, const double* d_t0,
                                   const double* d_t_stop, const double* d_dt0, double* d_y, const double* d_params,
                                   int32_t* d_status, double* d_dt_out, double rtol, double atol, double dt_min,
                                   int32_t grid, void* stream) {
  const odv::NetEntry* net = find_net(network);
  if (!net) return fail(ODEVEC_ENOTFOUND, std::string("unknown network ") + (network ? network : "(null)"));
  if (!d_t_stop || !d_y) return fail(ODEVEC_EINVAL, "NULL argument");
  if (precision != ODEVEC_FP64 && precision != ODEVEC_MIXED) return fail(ODEVEC_EINVAL, "bad precision");
  if (ncells < 0 || ncells > 0x7fffffffLL) return fail(ODEVEC_EINVAL, "ncells must be in [0, 2^31)");
  if (grid < 0) return fail(ODEVEC_EINVAL, "grid must be >= 0");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return fail(ODEVEC_EINVAL, "need rtol > 0 and atol >= 0");
  if (ncells == 0) return ODEVEC_OK;
  const int mode = net->lane_ok ? ODEVEC_PERSISTENT : (net->coop_ok ? ODEVEC_COOP : -1);
  if (mode < 0) return fail(ODEVEC_EINVAL, "network has no persistent kernel");
  if (!d_params && net->nparam > NOALLOC_MAXP)
    return fail(ODEVEC_EINVAL, "network has more parameters than the scratch holds defaults for; pass d_params");
  const size_t ws_cap = (ODEVEC_NOALLOC_SCRATCH_BYTES - NOALLOC_HEAD) / 8;   // doubles
  odv::Plan p{};
  hipError_t e = net->plan(ncells, mode, precision, 0, grid, 1, &p);
  if (e != hipSuccess) return hip_fail(e, "launch plan");
  if ((size_t)p.ws_doubles > ws_cap) {
    const long long per_wg = p.ws_doubles / p.grid;
    if (grid > 0 || per_wg <= 0 || (long long)ws_cap < per_wg) {
      char buf[200];
      std::snprintf(buf, sizeof buf, "%lld workgroups need %lld bytes of workspace; the scratch holds %zu",
                    (long long)p.grid, p.ws_doubles * 8, ws_cap * 8);
      return fail(ODEVEC_EINVAL, buf);
    }
    // the full-device grid does not fit: the largest grid that does
    e = net->plan(ncells, mode, precision, 0, (int)((long long)ws_cap / per_wg), 1, &p);
    if (e != hipSuccess) return hip_fail(e, "launch plan");
  }
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc != ODEVEC_OK) return rc;
  unsigned char* b = static_cast<unsigned char*>(base);
  double* defaults = reinterpret_cast<double*>(b + 256);
  hipStream_t st = (hipStream_t)stream;
  odv::KArgs a{};
  a.y = d_y;
  a.ncells = ncells;
  a.tstop = d_t_stop;
  a.ts_s = 1;
  if (d_t0) { a.t0 = d_t0; a.t0_s = 1; }
  else {
    hipLaunchKernelGGL(odv::k_fill_f64, dim3(1), dim3(64), 0, st, defaults, 0.0, 1LL);
    a.t0 = defaults; a.t0_s = 0;
  }
  if (d_dt0) { a.dt0 = d_dt0; a.dt_s = 1; }
  else {
    hipLaunchKernelGGL(odv::k_fill_f64, dim3(1), dim3(64), 0, st, defaults + 1, 1e-10, 1LL);
    a.dt0 = defaults + 1; a.dt_s = 0;
  }
  if (d_params && net->nparam > 0) {
    a.prm = d_params; a.ps_cell = 1; a.ps_plane = ncells;
  } else {
    double pd[NOALLOC_MAXP > 0 ? NOALLOC_MAXP : 1] = {};
    if (net->nparam > 0) net->param_defaults(pd);
    for (int k = 0; k < net->nparam; ++k)
      hipLaunchKernelGGL(odv::k_fill_f64, dim3(1), dim3(64), 0, st, defaults + 2 + k, pd[k], 1LL);
    a.prm = defaults + 2; a.ps_cell = 0; a.ps_plane = 1;
  }
  a.status = d_status;
  a.dt_out = d_dt_out;
  a.queue = reinterpret_cast<unsigned long long*>(b);
  a.ws = reinterpret_cast<double*>(b + NOALLOC_HEAD);
  a.ctl.rtol = rtol;
  a.ctl.atol = atol;
  a.ctl.dt_min = dt_min;
  a.ctl.max_attempts = 1000000;
  a.ctl.inv_tol = ODEVEC_MIXED_INVARIANT_TOL;
  a.ctl.promote_inv = ODEVEC_FALLBACK_INVARIANT_FAILS;
  a.group = 1;
  e = net->launch(a, mode, precision, p, st);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return ODEVEC_OK;
}

}  // extern "C"

/* ------------------------------------------------------------------ batches */

namespace {

const odv::BundleEntry* find_bundle(const char* id) {
  if (!id) return nullptr;
  std::string s(id);
  int repro = 0;
  const std::string suf = "+REPRO";
  if (s.size() > suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0) {
    repro = 1;
    s.resize(s.size() - suf.size());
  }
  for (const odv::BundleEntry* e : odv::bundles())
    if (e->repro == repro && s == e->id) return e;
  return nullptr;
}

constexpr int MAXP = 8;   // parameter-default slots per entry

}  // namespace

struct odevec_batch_s {
  const odv::BundleEntry* b = nullptr;
  int block = 0, grid = 0, use_graph = 0;
  odv::SegDev* h_segs = nullptr;   // pinned staging of the segment table
  int* h_nseg = nullptr;
  odv::SegDev* d_segs = nullptr;
  int* d_nseg = nullptr;
  unsigned long long* d_queues = nullptr;
  double* d_scal = nullptr;        // [0] t0 default 0, [1] dt hint default 1e-10, [2..] per-entry param defaults
  double* d_ws = nullptr;
  int64_t ws_cap = 0;
  hipEvent_t ev_copied = nullptr;  // the staging table has been copied (it may be rewritten)
  bool ev_pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  odv::Plan gplan{};
  hipStream_t gstream = nullptr;
  int ghist = 0;                   // the captured launch is the history variant
  int promote_cf = 0;              // FP64 fallback of mixed segments (0: off)
  double inv_tol = ODEVEC_MIXED_INVARIANT_TOL;   // mixed segments: invariant test of the corrector (0: off)
  int promote_inv = ODEVEC_FALLBACK_INVARIANT_FAILS;   // after this many failed tests: FP64 fallback, or no test (0: neither)
  int* d_promo = nullptr;          // handed-back cells of every mixed segment, one region per segment
  int64_t promo_cap = 0;
  double* d_ho = nullptr;          // their hand-off records, one region per segment
  int64_t ho_cap = 0;
  unsigned* d_done = nullptr;      // MAXSEG counts of batch wavefronts past each segment
  unsigned long long* d_cq = nullptr;   // MAXSEG claim counters of the FP64 lanes
  hipStream_t fb_stream = nullptr;      // the FP64 lanes run beside the batch launch on this stream
  hipEvent_t fb_fork = nullptr, fb_join = nullptr;
  unsigned* d_promo_n = nullptr;   // MAXSEG counts
  double* d_ws2 = nullptr;         // workspace of the FP64 passes
  int64_t ws2_cap = 0;
  int* d_order = nullptr;          // per-lane segments larger than the grid: queue in cost order
  unsigned char* d_bin = nullptr;
  int64_t order_cap = 0;
  unsigned* d_count = nullptr;     // MAXSEG x COST_BINS
};

namespace {

void batch_drop_graph(odevec_batch_s* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// The FP64 passes over the cells the batch launch handed back: one launch
// of the network's FP64 persistent kernel per mixed segment, its queue the
// segment's list (order) and device count (qlen), on the batch's stream.
int batch_fallback_passes(odevec_batch_s* b, const odevec_segment_t* segs, int nseg,
                          const std::vector<const odv::NetEntry*>& fbnet, hipStream_t st) {
  for (int k = 0; k < nseg; ++k) {
    if (!fbnet[k]) continue;
    odv::Plan p2{};
    hipError_t he = fbnet[k]->plan(segs[k].ncells, ODEVEC_PERSISTENT, ODEVEC_FP64, b->block, 0, 1, &p2);
    if (he != hipSuccess) return hip_fail(he, "launch plan (FP64 pass)");
    if (p2.ws_doubles > b->ws2_cap) {
      if (b->d_ws2) (void)hipFree(b->d_ws2);
      b->d_ws2 = nullptr;
      b->ws2_cap = 0;
      if ((he = hipMalloc((void**)&b->d_ws2, (size_t)p2.ws_doubles * 8)) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback workspace allocation failed: ") + hipGetErrorString(he));
      b->ws2_cap = p2.ws_doubles;
    }
    odv::KArgs a2 = b->h_segs[k].a;
    a2.ctl.promote_cf = 0;
    // handed-back cells run to the end (-1: a quantum segment's, whose
    // history holds pause records only)
    a2.quantum = a2.quantum > 0 ? -1 : 0;
    a2.order = a2.promo;
    a2.qlen = a2.promo_n;
    a2.ho_mode = 2;
    a2.fb_done = nullptr;
    a2.promo = nullptr;
    a2.promo_n = nullptr;
    a2.ws = b->d_ws2;
    he = fbnet[k]->launch(a2, ODEVEC_PERSISTENT, ODEVEC_FP64, p2, st);
    if (he != hipSuccess) return hip_fail(he, "kernel launch (FP64 pass)");
  }
  return ODEVEC_OK;
}

// The FP64 lanes beside a batch launch: one odv_k_fb_consume per mixed
// segment with the fallback, in table order on the batch's second stream
// (each ends once every batch wavefront has passed its segment and its
// handed-back cells are done), joined back into the batch's stream; then
// the FP64 passes over what they left.
int batch_fallback_lanes(odevec_batch_s* b, const odevec_segment_t* segs, int nseg,
                         const std::vector<const odv::NetEntry*>& fbnet, int nprod, long long nfb, hipStream_t st) {
  hipError_t he;
  for (int k = 0; k < nseg; ++k) {
    if (!fbnet[k]) continue;
    odv::KArgs c = b->h_segs[k].a;
    c.ctl.promote_cf = 0;
    c.quantum = c.quantum > 0 ? -1 : 0;
    c.order = c.promo;
    c.qlen = c.promo_n;
    c.promo = nullptr;
    c.promo_n = nullptr;
    c.fb_done = nullptr;
    c.ws = b->d_ws2;
    c.ho_mode = 2;
    c.queue = b->d_cq + k;
    he = fbnet[k]->launch_fbc(c, b->d_done + k, nprod, nfb, b->fb_stream);
    if (he != hipSuccess) return hip_fail(he, "kernel launch (FP64 lanes)");
  }
  if ((he = hipEventRecord(b->fb_join, b->fb_stream)) != hipSuccess ||
      (he = hipStreamWaitEvent(st, b->fb_join, 0)) != hipSuccess)
    return hip_fail(he, "fallback join");
  return batch_fallback_passes(b, segs, nseg, fbnet, st);
}

}  // namespace

int32_t odevec_bundle_count(void) { return (int32_t)odv::bundles().size(); }

const char* odevec_bundle_name(int32_t i) {
  static std::vector<std::string> names;
  if (i < 0 || i >= (int32_t)odv::bundles().size()) return nullptr;
  if (names.size() != odv::bundles().size()) {
    names.clear();
    for (const odv::BundleEntry* e : odv::bundles()) names.push_back(std::string(e->id) + (e->repro ? "+REPRO" : ""));
  }
  return names[(size_t)i].c_str();
}

int32_t odevec_bundle_entry(const char* bundle, int32_t index, const char** network, int32_t* precision) {
  const odv::BundleEntry* e = find_bundle(bundle);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown bundle ") + (bundle ? bundle : "(null)"));
  if (index < 0 || index >= e->n) return fail(ODEVEC_EINVAL, "entry index out of range");
  if (network) *network = e->nets[index];
  if (precision) *precision = e->precs[index];
  return e->n;
}

int32_t odevec_batch_create(const char* bundle, int32_t block, int32_t grid, int32_t use_graph, odevec_batch_t** out) {
  if (!out) return fail(ODEVEC_EINVAL, "out is NULL");
  *out = nullptr;
  const odv::BundleEntry* e = find_bundle(bundle);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown bundle ") + (bundle ? bundle : "(null)"));
  if (block != 0 && (block < 64 || block > 256 || block % 64)) return fail(ODEVEC_EINVAL, "block must be 64..256, a multiple of 64");
  if (grid < 0) return fail(ODEVEC_EINVAL, "grid must be >= 0");
  if (e->segdev_bytes != (int)sizeof(odv::SegDev)) return fail(ODEVEC_EINVAL, "bundle built with another segment layout");
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev < 1) return fail(ODEVEC_EDEVICE, "no HIP device available");
  odevec_batch_s* b = new odevec_batch_s();
  b->b = e;
  b->block = block;
  b->grid = grid;
  b->use_graph = use_graph ? 1 : 0;
  const size_t tbytes = sizeof(odv::SegDev) * odv::MAXSEG;
  const int nsc = 2 + e->n * MAXP;
  if ((he = hipHostMalloc((void**)&b->h_segs, tbytes, hipHostMallocDefault)) != hipSuccess ||
      (he = hipHostMalloc((void**)&b->h_nseg, sizeof(int), hipHostMallocDefault)) != hipSuccess ||
      (he = hipMalloc((void**)&b->d_segs, tbytes)) != hipSuccess ||
      (he = hipMalloc((void**)&b->d_nseg, sizeof(int))) != hipSuccess ||
      (he = hipMalloc((void**)&b->d_queues, odv::MAXSEG * sizeof(unsigned long long))) != hipSuccess ||
      (he = hipMalloc((void**)&b->d_scal, nsc * sizeof(double))) != hipSuccess ||
      (he = hipEventCreateWithFlags(&b->ev_copied, hipEventDisableTiming)) != hipSuccess) {
    odevec_batch_destroy(b);
    return fail(ODEVEC_ENOMEM, std::string("batch allocation failed: ") + hipGetErrorString(he));
  }
  std::vector<double> sc(nsc, 0.0);
  sc[1] = 1e-10;
  for (int k = 0; k < e->n; ++k) {
    if (e->nparam[k] > MAXP) {
      odevec_batch_destroy(b);
      return fail(ODEVEC_EINVAL, "network with more than 8 parameters in a bundle");
    }
    if (e->nparam[k] > 0) e->param_defaults[k](sc.data() + 2 + k * MAXP);
  }
  if ((he = hipMemcpy(b->d_scal, sc.data(), nsc * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) {
    odevec_batch_destroy(b);
    return hip_fail(he, "batch defaults upload");
  }
  *out = b;
  return ODEVEC_OK;
}

int32_t odevec_batch_solve_async(odevec_batch_t* b, const odevec_segment_t* segs, int32_t nseg, void* stream) {
  if (!b || (!segs && nseg > 0)) return fail(ODEVEC_EINVAL, "NULL argument");
  if (nseg < 0 || nseg > odv::MAXSEG) return fail(ODEVEC_EINVAL, "nseg must be 0..1024");
  const odv::BundleEntry* e = b->b;
  hipStream_t st = (hipStream_t)stream;
  // map and check every segment before touching the staging table
  std::vector<int> kind(nseg);
  int64_t total = 0;
  int hist = 0;
  for (int k = 0; k < nseg; ++k) {
    const odevec_segment_t& g = segs[k];
    int kk = -1;
    for (int j = 0; j < e->n; ++j)
      if (g.network && !std::strcmp(g.network, e->nets[j]) && g.precision == e->precs[j]) kk = j;
    if (kk < 0)
      return fail(ODEVEC_ENOTFOUND, std::string("segment network/precision not in bundle ") + e->id + ": " +
                                        (g.network ? g.network : "(null)"));
    if (g.ncells < 0) return fail(ODEVEC_EINVAL, "segment ncells < 0");
    if (g.ncells > 0 && (!g.y || !g.t_stop)) return fail(ODEVEC_EINVAL, "segment y / t_stop is NULL");
    if (!(g.rtol > 0.0) || !(g.atol >= 0.0)) return fail(ODEVEC_EINVAL, "segment tolerances");
    if (g.quantum < 0 || (g.quantum > 0 && (!g.hist || !g.status || g.flops)))
      return fail(ODEVEC_EINVAL, "segment quantum: > 0 needs hist and status, and no flops");
    kind[k] = kk;
    total += g.ncells;
    if (g.hist && hist < 1) hist = 1;
    if (g.quantum > 0) hist = 2;
  }
  if (total == 0) return ODEVEC_OK;
  // FP64 fallback: mixed segments hand back the cells that reach
  // promote_cf convergence failures; FP64 lanes beside the batch launch (one
  // kernel per such segment, on a second stream, in 1 in FB_SHARE_BATCH of the
  // machine's wavefront slots) continue them from their hand-off records, and
  // after both the network's FP64 persistent kernel takes what they left
  std::vector<const odv::NetEntry*> fbnet(nseg, nullptr);
  std::vector<int64_t> fboff(nseg, 0), hooff(nseg, 0);
  int64_t fbtotal = 0, hototal = 0;
  if (b->promote_cf > 0) {
    for (int k = 0; k < nseg; ++k) {
      const odevec_segment_t& g = segs[k];
      if (g.precision != ODEVEC_MIXED || g.ncells <= 0) continue;
      if (g.ncells > 0x7fffffffLL)
        return fail(ODEVEC_EINVAL, "the FP64 fallback takes at most 2^31 - 1 cells per segment");
      const odv::NetEntry* ne = find_net_build(g.network, b->b->repro);
      if (!ne || !ne->lane_ok) continue;
      fbnet[k] = ne;
      fboff[k] = fbtotal;
      fbtotal += g.ncells;
      hooff[k] = hototal;
      hototal += g.ncells * (int64_t)ne->handoff_doubles;
    }
  }
  // a graph is captured for one launch shape: in graph mode the grid does not
  // depend on the batch's size (workgroups without cells leave at once), so
  // one captured launch serves every batch
  odv::Plan p{};
  hipError_t he = e->plan(b->use_graph ? (int64_t)1 << 40 : total, b->block, b->grid, &p);
  if (he != hipSuccess) return hip_fail(he, "batch plan");
  long long nfb = 0;
  int nprod = 0;
  if (fbtotal > 0) {
    const long long wpb = p.block / 64, waves = p.grid * wpb;
    nfb = waves / odv::FB_SHARE_BATCH > 0 ? waves / odv::FB_SHARE_BATCH : 1;
    const long long keep = (waves - nfb) / wpb;
    p.grid = keep > 0 ? keep : 1;
    nprod = (int)(p.grid * wpb);
  }
  if (p.ws_doubles > b->ws_cap) {
    batch_drop_graph(b);   // the captured launch holds the old workspace
    if (b->d_ws) (void)hipFree(b->d_ws);
    b->d_ws = nullptr;
    b->ws_cap = 0;
    if ((he = hipMalloc((void**)&b->d_ws, (size_t)p.ws_doubles * 8)) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("rate workspace allocation failed: ") + hipGetErrorString(he));
    b->ws_cap = p.ws_doubles;
  }
  if (fbtotal > 0) {
    if (fbtotal > b->promo_cap) {
      if (b->d_promo) (void)hipFree(b->d_promo);
      b->d_promo = nullptr;
      b->promo_cap = 0;
      if ((he = hipMalloc((void**)&b->d_promo, (size_t)fbtotal * sizeof(int))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback list allocation failed: ") + hipGetErrorString(he));
      b->promo_cap = fbtotal;
    }
    if (fbtotal > 0 && !b->d_promo_n &&
        (he = hipMalloc((void**)&b->d_promo_n, odv::MAXSEG * sizeof(unsigned))) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("fallback list allocation failed: ") + hipGetErrorString(he));
    if (hototal > b->ho_cap) {
      if (b->d_ho) (void)hipFree(b->d_ho);
      b->d_ho = nullptr;
      b->ho_cap = 0;
      if ((he = hipMalloc((void**)&b->d_ho, (size_t)hototal * sizeof(double))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback record allocation failed: ") + hipGetErrorString(he));
      b->ho_cap = hototal;
    }
    // the FP64 lanes beside the launch: their counters, stream and workspace
    // (one segment's kernel at a time on that stream)
    if (!b->d_done && (he = hipMalloc((void**)&b->d_done, odv::MAXSEG * sizeof(unsigned))) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("fallback counter allocation failed: ") + hipGetErrorString(he));
    if (!b->d_cq && (he = hipMalloc((void**)&b->d_cq, odv::MAXSEG * sizeof(unsigned long long))) != hipSuccess)
      return fail(ODEVEC_ENOMEM, std::string("fallback counter allocation failed: ") + hipGetErrorString(he));
    if (!b->fb_stream && (he = hipStreamCreateWithFlags(&b->fb_stream, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(he, "fallback stream");
    if (!b->fb_fork && (he = hipEventCreateWithFlags(&b->fb_fork, hipEventDisableTiming)) != hipSuccess)
      return hip_fail(he, "fallback event");
    if (!b->fb_join && (he = hipEventCreateWithFlags(&b->fb_join, hipEventDisableTiming)) != hipSuccess)
      return hip_fail(he, "fallback event");
    long long ws64 = 0;
    for (int k = 0; k < nseg; ++k)
      if (fbnet[k] && nfb * fbnet[k]->ws64_per_wave > ws64) ws64 = nfb * fbnet[k]->ws64_per_wave;
    if (ws64 > b->ws2_cap) {
      if (b->d_ws2) (void)hipFree(b->d_ws2);
      b->d_ws2 = nullptr;
      b->ws2_cap = 0;
      if ((he = hipMalloc((void**)&b->d_ws2, (size_t)ws64 * 8)) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("fallback workspace allocation failed: ") + hipGetErrorString(he));
      b->ws2_cap = ws64;
    }
  }
  // the previous call's table copy must have left the staging buffer
  if (b->ev_pending && (he = hipEventSynchronize(b->ev_copied)) != hipSuccess) return hip_fail(he, "batch staging");
  for (int k = 0; k < nseg; ++k) {
    const odevec_segment_t& g = segs[k];
    odv::SegDev& d = b->h_segs[k];
    std::memset(&d, 0, sizeof d);
    odv::KArgs& a = d.a;
    a.y = g.y;
    a.ncells = g.ncells;
    if (g.params && e->nparam[kind[k]] > 0) { a.prm = g.params; a.ps_cell = 1; a.ps_plane = g.ncells; }
    else { a.prm = b->d_scal + 2 + kind[k] * MAXP; a.ps_cell = 0; a.ps_plane = 1; }
    if (g.t0) { a.t0 = g.t0; a.t0_s = 1; } else { a.t0 = b->d_scal; a.t0_s = 0; }
    a.tstop = g.t_stop; a.ts_s = 1;
    if (g.dt0) { a.dt0 = g.dt0; a.dt_s = 1; } else { a.dt0 = b->d_scal + 1; a.dt_s = 0; }
    a.mask = nullptr;
    a.status = g.status;
    a.dt_out = g.dt_out;
    a.counters = g.counters;
    a.flops = g.flops;
    a.queue = b->d_queues + k;
    a.ws = b->d_ws;
    a.ctl.rtol = g.rtol;
    a.ctl.atol = g.atol;
    a.ctl.dt_min = g.dt_min;
    a.ctl.check_negatives = g.check_negatives;
    a.ctl.max_attempts = g.max_attempts > 0 ? g.max_attempts : 1000000;
    a.ctl.inv_tol = b->inv_tol;
    a.ctl.promote_inv = b->promote_inv;
    a.group = 1;
    a.hist = g.hist;
    a.quantum = g.quantum;
    if (fbnet[k]) {
      a.ctl.promote_cf = b->promote_cf;
      a.promo = b->d_promo + fboff[k];
      a.promo_n = b->d_promo_n + k;
      a.ho = b->d_ho + hooff[k];
      a.ho_mode = 1;
      a.fb_done = b->d_done + k;
    }
    d.kind = kind[k];
  }
  // per-lane segments with more cells than the grid has lanes: the queue in
  // cost order (costliest first, as odevec_opts_t.schedule = 1 does for a
  // handle); results do not depend on it
  {
    int64_t ototal = 0;
    std::vector<const odv::NetEntry*> onet(nseg, nullptr);
    for (int k = 0; k < nseg; ++k) {
      const odv::NetEntry* ne = find_net_build(segs[k].network, b->b->repro);
      if (ne && ne->lane_ok && segs[k].ncells > p.grid * p.block && !segs[k].hist) {
        if (segs[k].ncells > 0x7fffffffLL)
          return fail(ODEVEC_EINVAL, "the cost-ordered queue takes at most 2^31 - 1 cells per segment");
        onet[k] = ne;
        ototal += segs[k].ncells;
      }
    }
    if (ototal > 0) {
      if (ototal > b->order_cap) {
        if (b->d_order) (void)hipFree(b->d_order);
        if (b->d_bin) (void)hipFree(b->d_bin);
        b->d_order = nullptr;
        b->d_bin = nullptr;
        b->order_cap = 0;
        if ((he = hipMalloc((void**)&b->d_order, (size_t)ototal * sizeof(int))) != hipSuccess ||
            (he = hipMalloc((void**)&b->d_bin, (size_t)ototal)) != hipSuccess)
          return fail(ODEVEC_ENOMEM, std::string("schedule allocation failed: ") + hipGetErrorString(he));
        b->order_cap = ototal;
      }
      if (!b->d_count &&
          (he = hipMalloc((void**)&b->d_count, (size_t)odv::MAXSEG * odv::COST_BINS * sizeof(unsigned))) != hipSuccess)
        return fail(ODEVEC_ENOMEM, std::string("schedule allocation failed: ") + hipGetErrorString(he));
      int64_t off = 0;
      for (int k = 0; k < nseg; ++k) {
        if (!onet[k]) continue;
        odv::KArgs& a = b->h_segs[k].a;
        if ((he = onet[k]->order(a, b->d_bin + off, b->d_count + (size_t)k * odv::COST_BINS, b->d_order + off, st)) !=
            hipSuccess)
          return hip_fail(he, "cost order");
        a.order = b->d_order + off;
        off += segs[k].ncells;
      }
    }
  }
  *b->h_nseg = nseg;
  if ((he = hipMemcpyAsync(b->d_segs, b->h_segs, sizeof(odv::SegDev) * nseg, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (he = hipMemcpyAsync(b->d_nseg, b->h_nseg, sizeof(int), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (he = hipEventRecord(b->ev_copied, st)) != hipSuccess)
    return hip_fail(he, "batch table upload");
  b->ev_pending = true;
  if (fbtotal > 0) {
    // lists read -1 until written; counts, wavefronts past, claims start at 0
    hipLaunchKernelGGL(odv::k_fill_int, dim3((unsigned)((fbtotal + 255) / 256)), dim3(256), 0, st, b->d_promo,
                       (long long)fbtotal, -1);
    hipLaunchKernelGGL(odv::k_zero_u, dim3((nseg + 63) / 64), dim3(64), 0, st, b->d_promo_n, nseg);
    hipLaunchKernelGGL(odv::k_zero_u, dim3((nseg + 63) / 64), dim3(64), 0, st, b->d_done, nseg);
    for (int k = 0; k < nseg; ++k)
      if (fbnet[k]) hipLaunchKernelGGL(odv::k_setq, dim3(1), dim3(1), 0, st, b->d_cq + k, 0ull);
    if ((he = hipEventRecord(b->fb_fork, st)) != hipSuccess ||
        (he = hipStreamWaitEvent(b->fb_stream, b->fb_fork, 0)) != hipSuccess)
      return hip_fail(he, "fallback fork");
  }
  if (!b->use_graph) {
    he = e->launch(b->d_segs, b->d_nseg, odv::MAXSEG, p, hist, st);
    if (he != hipSuccess) return hip_fail(he, "batch launch");
    return fbtotal > 0 ? batch_fallback_lanes(b, segs, nseg, fbnet, nprod, nfb, st) : ODEVEC_OK;
  }
  // graph: the queue reset and the batch kernel, captured once per launch
  // shape and stream and replayed; the table is read from device memory, so
  // one graph serves every batch of that shape
  if (b->exec && (b->gplan.grid != p.grid || b->gplan.block != p.block || b->gstream != st || b->ghist != hist))
    batch_drop_graph(b);
  if (!b->exec) {
    if ((he = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hip_fail(he, "graph capture");
    hipError_t le = e->launch(b->d_segs, b->d_nseg, odv::MAXSEG, p, hist, st);
    he = hipStreamEndCapture(st, &b->graph);
    if (le != hipSuccess || he != hipSuccess) {
      batch_drop_graph(b);
      return hip_fail(le != hipSuccess ? le : he, "graph capture");
    }
    if ((he = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) {
      batch_drop_graph(b);
      return hip_fail(he, "graph instantiate");
    }
    b->gplan = p;
    b->gstream = st;
    b->ghist = hist;
  }
  if ((he = hipGraphLaunch(b->exec, st)) != hipSuccess) return hip_fail(he, "graph launch");
  return fbtotal > 0 ? batch_fallback_lanes(b, segs, nseg, fbnet, nprod, nfb, st) : ODEVEC_OK;
}

int32_t odevec_batch_set_fp64_fallback(odevec_batch_t* b, int32_t convfail_limit) {
  if (!b) return fail(ODEVEC_EINVAL, "batch is NULL");
  if (convfail_limit < 0) return fail(ODEVEC_EINVAL, "convfail_limit must be >= 0");
  b->promote_cf = convfail_limit;
  return ODEVEC_OK;
}

int32_t odevec_batch_set_invariant_test(odevec_batch_t* b, double tolerances, int32_t fallback_fails) {
  if (!b) return fail(ODEVEC_EINVAL, "batch is NULL");
  if (!(tolerances >= 0.0)) return fail(ODEVEC_EINVAL, "tolerances must be >= 0");
  if (fallback_fails < 0) return fail(ODEVEC_EINVAL, "fallback_fails must be >= 0");
  b->inv_tol = tolerances;
  b->promote_inv = fallback_fails;
  return ODEVEC_OK;
}

int32_t odevec_batch_kernel_info(const char* bundle, int32_t block, odevec_kernel_info_t* info) {
  const odv::BundleEntry* e = find_bundle(bundle);
  if (!e) return fail(ODEVEC_ENOTFOUND, std::string("unknown bundle ") + (bundle ? bundle : "(null)"));
  if (!info) return fail(ODEVEC_EINVAL, "info is NULL");
  odv::KernelInfo k{};
  hipError_t he = e->info(block, &k);
  if (he != hipSuccess) return hip_fail(he, "kernel info");
  info->vgprs = k.regs; info->spill_bytes = k.spill_bytes;
  info->blocks_per_cu = k.blocks_per_cu; info->block = k.block;
  return ODEVEC_OK;
}

int32_t odevec_batch_destroy(odevec_batch_t* b) {
  if (!b) return ODEVEC_OK;
  if (b->ev_pending) (void)hipEventSynchronize(b->ev_copied);
  (void)hipDeviceSynchronize();   // no launch of this batch may still read its buffers
  batch_drop_graph(b);
  if (b->h_segs) (void)hipHostFree(b->h_segs);
  if (b->h_nseg) (void)hipHostFree(b->h_nseg);
  void* ps[] = {b->d_segs, b->d_nseg, b->d_queues, b->d_scal, b->d_ws, b->d_promo, b->d_promo_n, b->d_ws2,
                b->d_order, b->d_bin, b->d_count, b->d_ho, b->d_done, b->d_cq};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (b->ev_copied) (void)hipEventDestroy(b->ev_copied);
  if (b->fb_stream) (void)hipStreamDestroy(b->fb_stream);
  if (b->fb_fork) (void)hipEventDestroy(b->fb_fork);
  if (b->fb_join) (void)hipEventDestroy(b->fb_join);
  delete b;
  return ODEVEC_OK;
}

// Test harness: the product's per-lane integrator (odevec_integrator.h and the
// generated network headers) compiled for the host, so tests can compare its
// control flow and arithmetic with the oracle on the CPU. Not part of the
// library; built by tests/conftest.py with --offload-host-only.
#include <cstring>

#include "../../include/odevec.h"

#include "odevec_integrator.h"
#include "generated/net_ROBER_NONE_DENSE.h"
#include "generated/net_ROBER_NONE_CSC.h"
#include "generated/net_OREGO_NONE_DENSE.h"
#include "generated/net_OREGO_NONE_CSC.h"
#include "generated/net_PRIMORDIAL_CMK_CSC.h"
#include "generated/net_PRIMORDIAL_NONE_DENSE.h"
#include "generated/net_PRIMORDIAL_FF_CSC.h"
#include "generated/net_ROBER_NONE_DENSE_LU3.h"
#include "generated/net_PRIMORDIAL_NONE_DENSE_LU3.h"
#include "generated/net_ROBER_NONE_DENSE_LU2.h"
#include "generated/net_OREGO_NONE_DENSE_LU2.h"
#include "generated/net_KROME_FF_CSC.h"

namespace {

// KS > 0 runs the lane with its rates, error weights, previous correction
// and (ZX) Nordsieck columns 2..MAXORD in external strided columns, as the
// device kernels keep them in LDS and in the rate workspace.
template <class Net, class M, bool RCP, int KS, int ZX = 0, bool WX = false>
void run(long n, const double* t0, const double* ts, const double* dt0, double* y, const double* prm,
         const odv::Controls& c, int* status, double* dtout, int* cnt) {
  constexpr int N = Net::NEQ;
  constexpr int NA = Net::NACT;
  constexpr int NP = Net::NPARAM > 0 ? Net::NPARAM : 1;
  constexpr int S = KS > 0 ? KS : 1;
  using L = odv::Lane<odv::Cfg<Net, M, RCP, KS, false, ZX, WX>>;
  static double kcol[Net::NRATE * S];
  static double ecol[NA * S];
  static double zcol[(odv::MAXORD - 1) * NA * S];
  static double wcol[NA * S];
  for (long i = 0; i < n; ++i) {
    L s;
    const int off = KS > 0 ? (int)(i % KS) : 0;
    s.kx = kcol + off;
    s.eox = ecol + off;
    s.zx = zcol + off;
    s.wx = wcol + off;
    odv::LaneGroup g;
    double yy[N], p[NP];
    for (int j = 0; j < N; ++j) yy[j] = y[j * n + i];
    for (int q = 0; q < NP; ++q) p[q] = Net::NPARAM > 0 ? prm[q * n + i] : 0.0;
    s.active = true;
    s.cnt = odv::Counters{0, 0, 0, 0, 0, 0, 0, 0};
    s.dt = dt0[i];
    bool done;
    if (!(t0[i] < ts[i])) {
      s.status = odv::ST_SKIPPED;
      done = true;
    } else {
      odv::odv_begin(s, yy, p, t0[i], ts[i], dt0[i], c, g);
      done = false;
    }
    while (!done) done = odv::odv_attempt(s, c, g);
    if (s.status != odv::ST_SKIPPED) {
      double ya[Net::NACT];
      odv::odv_result(s, ya);
      for (int j = 0; j < Net::NACT; ++j) yy[Net::ACT[j]] = ya[j];
    }
    for (int j = 0; j < N; ++j) y[j * n + i] = yy[j];
    status[i] = s.status;
    dtout[i] = s.dt;
    const odv::Counters& k = s.cnt;
    const int v[8] = {k.attempts, k.steps, k.newton, k.jac, k.errfail, k.convfail, k.stepsel, k.qsum};
    for (int m = 0; m < 8; ++m) cnt[m * n + i] = v[m];
  }
}

template <class Net>
void run_prec(int mixed, long n, const double* t0, const double* ts, const double* dt0, double* y,
              const double* prm, const odv::Controls& c, int* status, double* dtout, int* cnt) {
  if (mixed & 2) {
    // the per-precision placement of the device kernels: mixed moves Nordsieck
    // columns 2..4 out of the lane, FP64 columns 2..MAXORD and the weights
    if (mixed & 1) run<Net, float, true, 7, 3, false>(n, t0, ts, dt0, y, prm, c, status, dtout, cnt);
    else run<Net, double, false, 7, 4, true>(n, t0, ts, dt0, y, prm, c, status, dtout, cnt);
  } else {
    if (mixed & 1) run<Net, float, true, 0>(n, t0, ts, dt0, y, prm, c, status, dtout, cnt);
    else run<Net, double, false, 0>(n, t0, ts, dt0, y, prm, c, status, dtout, cnt);
  }
}

}  // namespace

extern "C" int lane_solve_cells(const char* id, int mixed, long n, const double* t0, const double* ts,
                                const double* dt0, double* y, const double* prm, double rtol, double atol,
                                double dt_min, int check_neg, int max_att, int* status, double* dtout,
                                int* cnt) {
  // the mixed form runs with the library's default invariant test
  odv::Controls c{rtol, atol, dt_min, check_neg, max_att, 0, (mixed & 1) ? ODEVEC_MIXED_INVARIANT_TOL : 0.0,
                  ODEVEC_FALLBACK_INVARIANT_FAILS};
#define ODV_CASE(NET) \
  if (!std::strcmp(id, NET::ID)) { run_prec<NET>(mixed, n, t0, ts, dt0, y, prm, c, status, dtout, cnt); return 0; }
  ODV_CASE(odv_net_ROBER_NONE_DENSE)
  ODV_CASE(odv_net_ROBER_NONE_CSC)
  ODV_CASE(odv_net_OREGO_NONE_DENSE)
  ODV_CASE(odv_net_OREGO_NONE_CSC)
  ODV_CASE(odv_net_PRIMORDIAL_CMK_CSC)
  ODV_CASE(odv_net_PRIMORDIAL_NONE_DENSE)
  ODV_CASE(odv_net_PRIMORDIAL_FF_CSC)
  ODV_CASE(odv_net_ROBER_NONE_DENSE_LU3)
  ODV_CASE(odv_net_PRIMORDIAL_NONE_DENSE_LU3)
  ODV_CASE(odv_net_ROBER_NONE_DENSE_LU2)
  ODV_CASE(odv_net_OREGO_NONE_DENSE_LU2)
  ODV_CASE(odv_net_KROME_FF_CSC)
#undef ODV_CASE
  return -1;
}

// generated network functions, host build, for unit tests of the emitted code
extern "C" int lane_netfun(const char* id, int what, const double* y, const double* k, double hb, double* out) {
#define ODV_CASE(NET)                                                      \
  if (!std::strcmp(id, NET::ID)) {                                         \
    if (what == 0) NET::rhs<1>(y, k, out);                                    \
    else if (what == 1) NET::rates(y, k, out);                             \
    else if (what == 2) NET::template jac<double, 1>(y, k, hb, out);          \
    else if (what == 3) { NET::template jac<double, 1>(y, k, hb, out); NET::template lu<double>(out); } \
    else if (what == 4) { double r[NET::NEQ]; for (int j = 0; j < NET::NEQ; ++j) r[j] = y[j];          \
                          NET::template solve<double, double>(k, r);                                   \
                          for (int j = 0; j < NET::NEQ; ++j) out[j] = r[j]; }                          \
    return 0;                                                              \
  }
  ODV_CASE(odv_net_ROBER_NONE_DENSE)
  ODV_CASE(odv_net_ROBER_NONE_CSC)
  ODV_CASE(odv_net_OREGO_NONE_DENSE)
  ODV_CASE(odv_net_OREGO_NONE_CSC)
  ODV_CASE(odv_net_PRIMORDIAL_CMK_CSC)
  ODV_CASE(odv_net_PRIMORDIAL_NONE_DENSE)
  ODV_CASE(odv_net_PRIMORDIAL_FF_CSC)
  ODV_CASE(odv_net_ROBER_NONE_DENSE_LU3)
  ODV_CASE(odv_net_PRIMORDIAL_NONE_DENSE_LU3)
  ODV_CASE(odv_net_ROBER_NONE_DENSE_LU2)
  ODV_CASE(odv_net_OREGO_NONE_DENSE_LU2)
  ODV_CASE(odv_net_KROME_FF_CSC)
#undef ODV_CASE
  return -1;
}

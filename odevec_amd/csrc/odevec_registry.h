// Registry of compiled network variants. Each generated instantiation unit
// registers one entry whose launch function dispatches the three kernel forms
// in the two precisions for that network.
#pragma once
#include <vector>

#include "odevec_coop.h"

namespace odv {

enum Mode : int { MODE_PERCELL = 0, MODE_LOCKSTEP = 1, MODE_PERSISTENT = 2, MODE_COOP = 3 };
enum Prec : int { PREC_FP64 = 0, PREC_MIXED = 1 };

struct KernelInfo {
  int regs;            // VGPRs per lane (arch + acc)
  int sgprs;
  int spill_bytes;     // private (scratch) bytes per lane
  int blocks_per_cu;   // occupancy at the chosen block size
  int block;
};

// Launch shape of one solve: threads per workgroup, workgroups, and the rate
// workspace it needs (doubles).
struct Plan {
  int block;
  long long grid;
  long long ws_doubles;
};

struct NetEntry {
  const char* id;
  int repro;   // built with -ffp-contract=off (bit-identical to the oracle)
  int neq, nnz, nrate, nparam, rate_kind, tgas;
  int flops_rhs, flops_jac, flops_lu, flops_solve, flops_lu_r, flops_solve_r;
  int lane_ok, coop_ok;   // per-lane modes (percell, lockstep, persistent) / cooperative mode compiled
  int hist_planes;        // warm-restart history planes per cell (per-lane modes; the cooperative
                          // mode's layout for a network only it runs)
  int handoff_doubles;    // FP64 fallback hand-off record per cell (per-lane modes)
  hipError_t (*plan)(long long ncells, int mode, int prec, int block, int grid, int group, Plan* out);
  hipError_t (*launch)(const KArgs& a, int mode, int prec, const Plan& p, hipStream_t st);
  hipError_t (*info)(int mode, int prec, int block, KernelInfo* out);
  void (*param_defaults)(double* p);
  // persistent mode: rank the cells by expected cost into `order` (scratch: bin[ncells], count[COST_BINS])
  hipError_t (*order)(const KArgs& a, unsigned char* bin, unsigned* count, int* order, hipStream_t st);
  // GetRHS (what 0) / GetJac (what 1, P = I - hb*J) for a.ncells cells into `out`
  hipError_t (*netfun)(int what, const KArgs& a, double hb, double* out, hipStream_t st);
  int jac_entries;        // entries per cell of the GetJac layout (0: per-lane code not built)
  // persistent mode with the FP64 fallback: FP64 lanes beside the queue
  // kernel (odv_k_fb_consume, `grid` one-wavefront workgroups) and the
  // workspace doubles per wavefront they need
  hipError_t (*launch_fbc)(const KArgs& b, unsigned* done, int nprod, long long grid, hipStream_t st);
  long long ws64_per_wave;
};

std::vector<const NetEntry*>& registry();

// A bundle: one batch kernel compiled for several (network, precision)
// entries, launched over a table of segments (odv_k_batch).
constexpr int MAXSEG = 1024;   // segments per batch launch

struct BundleEntry {
  const char* id;              // e.g. "MIX3"
  int repro;
  int n;                       // entries
  const char* const* nets;     // network id of each entry
  const int* precs;            // precision of each entry (Prec)
  const int* neq;              // species of each entry's network
  const int* nparam;
  void (*const* param_defaults)(double*);
  long long ws_per_slot;       // rate workspace doubles per workgroup slot (max over entries)
  int segdev_bytes;            // sizeof(SegDev) of the build
  hipError_t (*plan)(long long ncells, int block, int grid, Plan* out);
  hipError_t (*launch)(void* segs, const int* nseg, int maxseg, const Plan& p, int hist, hipStream_t st);
  hipError_t (*info)(int block, KernelInfo* out);
};

std::vector<const BundleEntry*>& bundles();

struct BundleRegistrar {
  explicit BundleRegistrar(const BundleEntry* e) { bundles().push_back(e); }
};

struct Registrar {
  explicit Registrar(const NetEntry* e) { registry().push_back(e); }
};

__global__ void k_setq(unsigned long long* q, unsigned long long v);

inline namespace ODV_VARIANT_NS {

template <class K>
inline hipError_t kinfo(K kern, int block, KernelInfo* out, size_t dyn_lds = 0) {
  hipFuncAttributes at;
  hipError_t e = hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kern));
  if (e != hipSuccess) return e;
  int nb = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kern), block, dyn_lds);
  if (e != hipSuccess) return e;
  out->regs = at.numRegs;
  out->sgprs = 0;
  out->spill_bytes = (int)at.localSizeBytes;
  out->blocks_per_cu = nb;
  out->block = block;
  return hipSuccess;
}

// block: threads per workgroup of the per-cell kernels (a multiple of 64 up
// to MAX_BLOCK; 0: BLOCK); grid: workgroups of the persistent kernel (0: one
// wave of resident workgroups, CUs x occupancy); group: lockstep group size.
template <class Net, class M, bool RCP>
inline hipError_t plan_t(long long ncells, int mode, int block, int grid, int group, Plan* p) {
  using P = LdsPolicy<Net, M>;
  if (mode == MODE_COOP) {
    if constexpr (!Net::COOP) {
      return hipErrorInvalidValue;
    } else {
      // persistent groups: one wave of resident workgroups (CUs x occupancy)
      constexpr int GPB = coop_block<Net, M>() / CoopWidth<Net::NACT>::G;
      long long g = grid;
      if (g <= 0) {
        int dev = 0, ncu = 0, nb = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
          e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, reinterpret_cast<const void*>(odv_k_coop<Net, M, RCP>), coop_block<Net, M>(),
              coop_lds_bytes<Net, M>());
        if (e != hipSuccess) return e;
        g = (long long)ncu * (nb > 0 ? nb : 1);
      }
      const long long need = (ncells + GPB - 1) / GPB;
      if (g > need) g = need;
      p->block = coop_block<Net, M>();
      p->grid = g;
      p->ws_doubles = 0;
      return hipSuccess;
    }
  }
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    const int b = block > 0 ? block : BLOCK;
    if (mode != MODE_LOCKSTEP && (b > MAX_BLOCK || b % WAVE != 0)) return hipErrorInvalidValue;
    if (mode == MODE_PERCELL) {
      p->block = b;
      p->grid = (ncells + b - 1) / b;
    } else if (mode == MODE_PERSISTENT) {
      long long g = grid;
      if (g <= 0) {
        int dev = 0, ncu = 0, nb = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
          e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, reinterpret_cast<const void*>(odv_k_persistent<Net, M, RCP>), b, lane_lds_bytes(P::WORDS, b));
        if (e != hipSuccess) return e;
        g = (long long)ncu * (nb > 0 ? nb : 1);
      }
      const long long need = (ncells + b - 1) / b;
      if (g > need) g = need;
      p->block = b;
      p->grid = g;
    } else {
      const int grp = group > 0 ? group : 1;
      p->block = ((grp + 63) / 64) * 64;
      p->grid = (ncells + grp - 1) / grp;
    }
    p->ws_doubles = (mode == MODE_LOCKSTEP) ? 0 : p->grid * (p->block / WAVE) * P::WS_PER_SLOT;
    return hipSuccess;
  }
}

template <class Net, class M, bool RCP>
inline hipError_t launch_t(const KArgs& a, int mode, const Plan& p, hipStream_t st) {
  if (a.ncells <= 0) return hipSuccess;
  if (p.block <= 0 || p.grid <= 0 || p.grid > 0x7fffffffLL) return hipErrorInvalidValue;
  if (mode != MODE_LOCKSTEP && (p.block > MAX_BLOCK || p.block % WAVE != 0)) return hipErrorInvalidValue;
  if (p.ws_doubles > 0 && !a.ws) return hipErrorInvalidValue;
  if (mode == MODE_COOP) {
    if constexpr (Net::COOP) {
      constexpr int CB = coop_block<Net, M>(), LDS = coop_lds_bytes<Net, M>();
      if (p.block != CB) return hipErrorInvalidValue;
      constexpr long long GPB = CB / CoopWidth<Net::NACT>::G;
      // a history in this mode only for networks no per-lane kernel runs
      // (a network has one history layout)
      if (a.hist && Net::LANE) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_setq, dim3(1), dim3(1), 0, st, a.queue, (unsigned long long)(p.grid * GPB));
      if (a.hist)
        hipLaunchKernelGGL((odv_k_coop<Net, M, RCP, 1>), dim3((unsigned)p.grid), dim3(CB), LDS, st, a);
      else
        hipLaunchKernelGGL((odv_k_coop<Net, M, RCP>), dim3((unsigned)p.grid), dim3(CB), LDS, st, a);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    const int lds = lane_lds_bytes(LdsPolicy<Net, M>::WORDS, p.block);
    if (a.hist && mode != MODE_PERCELL && mode != MODE_PERSISTENT) return hipErrorInvalidValue;
    if (mode == MODE_PERCELL) {
      if (a.hist)
        hipLaunchKernelGGL((odv_k_percell<Net, M, RCP, 1>), dim3((unsigned)p.grid), dim3(p.block), lds, st, a);
      else
        hipLaunchKernelGGL((odv_k_percell<Net, M, RCP>), dim3((unsigned)p.grid), dim3(p.block), lds, st, a);
    } else if (mode == MODE_PERSISTENT) {
      hipLaunchKernelGGL(k_setq, dim3(1), dim3(1), 0, st, a.queue, (unsigned long long)(p.grid * p.block));
      if (a.hist)
        hipLaunchKernelGGL((odv_k_persistent<Net, M, RCP, 1>), dim3((unsigned)p.grid), dim3(p.block), lds, st,
                           a);
      else
        hipLaunchKernelGGL((odv_k_persistent<Net, M, RCP>), dim3((unsigned)p.grid), dim3(p.block), lds, st, a);
    } else {
      hipLaunchKernelGGL((odv_k_lockstep<Net, M, RCP>), dim3((unsigned)p.grid), dim3(p.block), 0, st, a);
    }
    return hipGetLastError();
  }
}

// FP64 lanes of the fallback beside a persistent launch: 1 in FB_SHARE of
// the machine's wavefront slots (at least one).
constexpr int FB_SHARE = 32;
// ... and beside a batch launch, 1 in FB_SHARE_BATCH (2,097,152-cell MIX3_MIXED
// batch, ms per solve at 1 in 8 / 16 / 32: 117.6 / 110.0 / 119.5; the 1M-cell
// PRIMORDIAL handle at 1 in 16 / 32: 119.5 / 115.7)
constexpr int FB_SHARE_BATCH = 16;
// an FP64 wavefront that has looked at the list this many times in a row with
// nothing to run (a sleep and a few loads each, about 1 us: about a second)
// leaves the rest of the fallback to the pass after the launch
constexpr long long FB_IDLE_ROUNDS = 1 << 20;

template <class Net>
hipError_t launch_fbc_net(const KArgs& b, unsigned* done, int nprod, long long grid, hipStream_t st) {
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    if (grid <= 0 || grid > 0x7fffffffLL || nprod <= 0) return hipErrorInvalidValue;
    if ((LdsPolicy<Net, double>::WS_PER_SLOT > 0 && !b.ws) || !b.order || !b.qlen || !b.queue || !b.ho || !done)
      return hipErrorInvalidValue;
    if (b.hist)
      hipLaunchKernelGGL((odv_k_fb_consume<Net, 1>), dim3((unsigned)grid), dim3(WAVE),
                         lane_lds_bytes(LdsPolicy<Net, double>::WORDS, WAVE), st, b, done, (unsigned)nprod,
                         FB_IDLE_ROUNDS);
    else
      hipLaunchKernelGGL((odv_k_fb_consume<Net, 0>), dim3((unsigned)grid), dim3(WAVE),
                         lane_lds_bytes(LdsPolicy<Net, double>::WORDS, WAVE), st, b, done, (unsigned)nprod,
                         FB_IDLE_ROUNDS);
    return hipGetLastError();
  }
}

// the per-lane kernels' history layout, or the cooperative kernel's for a
// network only it runs
template <class Net>
constexpr int hist_planes_of() {
  if constexpr (Net::LANE || !Net::COOP) return Hist<Net>::PLANES;
  else return CoopHist<Net>::PLANES;
}

template <class Net>
constexpr long long ws64_per_wave_of() {
  if constexpr (Net::LANE) return LdsPolicy<Net, double>::WS_PER_SLOT;
  else return 0;
}

template <class Net>
hipError_t netfun_net(int what, const KArgs& a, double hb, double* out, hipStream_t st) {
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    const unsigned g = (unsigned)((a.ncells + 63) / 64);
    if (what == 0) hipLaunchKernelGGL(k_get_rhs<Net>, dim3(g), dim3(64), 0, st, a, out);
    else hipLaunchKernelGGL(k_get_jac<Net>, dim3(g), dim3(64), 0, st, a, hb, out);
    return hipGetLastError();
  }
}

template <class Net>
constexpr int jac_entries_of() {
  if constexpr (Net::LANE) return Net::JMAP_N;
  else return 0;
}

template <class Net>
hipError_t order_net(const KArgs& a, unsigned char* bin, unsigned* count, int* order, hipStream_t st) {
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    const unsigned g = (unsigned)((a.ncells + 255) / 256);
    hipLaunchKernelGGL(k_cost_zero<0>, dim3(1), dim3(COST_BINS), 0, st, count);
    hipLaunchKernelGGL(k_cost_bin<Net>, dim3(g), dim3(256), 0, st, a, bin, count);
    hipLaunchKernelGGL(k_cost_scan<0>, dim3(1), dim3(COST_BINS), 0, st, count);
    hipLaunchKernelGGL(k_cost_scatter<0>, dim3(g), dim3(256), 0, st, (long long)a.ncells,
                       static_cast<const unsigned char*>(bin), count, order);
    return hipGetLastError();
  }
}

template <class Net>
hipError_t plan_net(long long ncells, int mode, int prec, int block, int grid, int group, Plan* p) {
  if (prec == PREC_MIXED) return plan_t<Net, float, true>(ncells, mode, block, grid, group, p);
  return plan_t<Net, double, false>(ncells, mode, block, grid, group, p);
}

template <class Net>
hipError_t launch_net(const KArgs& a, int mode, int prec, const Plan& p, hipStream_t st) {
  if (prec == PREC_MIXED) return launch_t<Net, float, true>(a, mode, p, st);
  return launch_t<Net, double, false>(a, mode, p, st);
}

template <class Net, class M, bool RCP>
inline hipError_t info_t(int mode, int block, KernelInfo* out) {
  if (mode == MODE_COOP) {
    if constexpr (Net::COOP)
      return kinfo(odv_k_coop<Net, M, RCP>, coop_block<Net, M>(), out, coop_lds_bytes<Net, M>());
    return hipErrorInvalidValue;
  }
  if constexpr (!Net::LANE) {
    return hipErrorInvalidValue;
  } else {
    const int b = block > 0 ? block : BLOCK;
    if (mode != MODE_LOCKSTEP && (b > MAX_BLOCK || b % WAVE != 0)) return hipErrorInvalidValue;
    const int lds = lane_lds_bytes(LdsPolicy<Net, M>::WORDS, b);
    if (mode == MODE_PERCELL) return kinfo(odv_k_percell<Net, M, RCP>, b, out, lds);
    if (mode == MODE_PERSISTENT) return kinfo(odv_k_persistent<Net, M, RCP>, b, out, lds);
    return kinfo(odv_k_lockstep<Net, M, RCP>, block > 0 ? block : 64, out);
  }
}

template <class Net>
hipError_t info_net(int mode, int prec, int block, KernelInfo* out) {
  if (prec == PREC_MIXED) return info_t<Net, float, true>(mode, block, out);
  return info_t<Net, double, false>(mode, block, out);
}

// Plan, launch and kernel info of one bundle's batch kernel.
template <class... E>
struct Bundle {
  static constexpr long long WS = MaxWords<E...>::WS;

  static hipError_t plan(long long ncells, int block, int grid, Plan* p) {
    const int b = block > 0 ? block : BLOCK;
    if (b > MAX_BLOCK || b % WAVE != 0) return hipErrorInvalidValue;
    long long g = grid;
    if (g <= 0) {
      int dev = 0, ncu = 0, nb = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(odv_k_batch<0, E...>), b,
                                                         lane_lds_bytes(MaxWords<E...>::WORDS, b));
      if (e != hipSuccess) return e;
      g = (long long)ncu * (nb > 0 ? nb : 1);
    }
    const long long need = (ncells + b - 1) / b;
    if (g > need) g = need > 0 ? need : 1;
    p->block = b;
    p->grid = g;
    p->ws_doubles = g * (b / WAVE) * WS;
    return hipSuccess;
  }

  // hist: 1 launch the variant that reads and writes the segments'
  // warm-restart histories, 2 the one that also runs launch quanta
  static hipError_t launch(void* segs, const int* nseg, int maxseg, const Plan& p, int hist, hipStream_t st) {
    if (p.block <= 0 || p.block > MAX_BLOCK || p.block % WAVE != 0 || p.grid <= 0 || p.grid > 0x7fffffffLL)
      return hipErrorInvalidValue;
    SegDev* sd = static_cast<SegDev*>(segs);
    hipLaunchKernelGGL(k_batch_reset<0>, dim3((unsigned)((maxseg + 255) / 256)), dim3(256), 0, st, sd, nseg,
                       (unsigned long long)(p.grid * p.block));
    if (hist == 2)
      hipLaunchKernelGGL((odv_k_batch<2, E...>), dim3((unsigned)p.grid), dim3(p.block),
                         lane_lds_bytes(MaxWords<E...>::WORDS, p.block), st, static_cast<const SegDev*>(sd), nseg);
    else if (hist)
      hipLaunchKernelGGL((odv_k_batch<1, E...>), dim3((unsigned)p.grid), dim3(p.block),
                         lane_lds_bytes(MaxWords<E...>::WORDS, p.block), st, static_cast<const SegDev*>(sd), nseg);
    else
      hipLaunchKernelGGL((odv_k_batch<0, E...>), dim3((unsigned)p.grid), dim3(p.block),
                         lane_lds_bytes(MaxWords<E...>::WORDS, p.block), st, static_cast<const SegDev*>(sd), nseg);
    return hipGetLastError();
  }

  static hipError_t info(int block, KernelInfo* out) {
    const int b = block > 0 ? block : BLOCK;
    if (b > MAX_BLOCK || b % WAVE != 0) return hipErrorInvalidValue;
    return kinfo(odv_k_batch<0, E...>, b, out, lane_lds_bytes(MaxWords<E...>::WORDS, b));
  }
};

}  // namespace ODV_VARIANT_NS
}  // namespace odv

// Register a bundle of (network, M, RCP) entries; NETS/PRECS/NEQ/NPARAM/DEFAULTS
// are arrays of the entries' network ids, precisions, species counts,
// parameter counts and parameter-default functions, in entry order.
#define ODV_REGISTER_BUNDLE(NAME, BUNDLE, NETS, PRECS, NEQ, NPARAM, DEFAULTS)                               \
  static const odv::BundleEntry odv_bundle_##NAME = {#NAME, ODV_REPRO_FLAG, (int)(sizeof(PRECS) / sizeof(int)), \
                                                     NETS, PRECS, NEQ, NPARAM, DEFAULTS, BUNDLE::WS,            \
                                                     (int)sizeof(odv::SegDev), &BUNDLE::plan, &BUNDLE::launch,   \
                                                     &BUNDLE::info};                                             \
  static odv::BundleRegistrar odv_bundle_reg_##NAME(&odv_bundle_##NAME);

#define ODV_REGISTER(NET)                                                                       \
  static const odv::NetEntry odv_entry_##NET = {  /* NOLINT */                                                \
      NET::ID, ODV_REPRO_FLAG, NET::NEQ, NET::NNZ, NET::NRATE, NET::NPARAM, NET::RATE_KIND, NET::TGAS,          \
      NET::FLOPS_RHS, NET::FLOPS_JAC, NET::FLOPS_LU, NET::FLOPS_SOLVE, NET::FLOPS_LU_R,         \
      NET::FLOPS_SOLVE_R, NET::LANE ? 1 : 0, NET::COOP ? 1 : 0, odv::hist_planes_of<NET>(), odv::Handoff<NET>::REC, &odv::plan_net<NET>, &odv::launch_net<NET>, &odv::info_net<NET>, &NET::param_defaults, &odv::order_net<NET>, &odv::netfun_net<NET>, odv::jac_entries_of<NET>(), &odv::launch_fbc_net<NET>, odv::ws64_per_wave_of<NET>()};                          \
  static odv::Registrar odv_reg_##NET(&odv_entry_##NET);

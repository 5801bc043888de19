// Registry of compiled network variants. Each generated instantiation unit
// registers one entry whose launch function dispatches the three kernel forms
// in the two precisions for that network.
#pragma once
#include <vector>

#include "odevec_kernels.h"

namespace odv {

enum Mode : int { MODE_PERCELL = 0, MODE_LOCKSTEP = 1, MODE_PERSISTENT = 2 };
enum Prec : int { PREC_FP64 = 0, PREC_MIXED = 1 };

struct KernelInfo {
  int regs;            // VGPRs per lane (arch + acc)
  int sgprs;
  int spill_bytes;     // private (scratch) bytes per lane
  int blocks_per_cu;   // occupancy at the chosen block size
  int block;
};

struct NetEntry {
  const char* id;
  int repro;   // built with -ffp-contract=off (bit-identical to the oracle)
  int neq, nnz, nrate, nparam, rate_kind, tgas;
  int flops_rhs, flops_jac, flops_lu, flops_solve, flops_lu_r, flops_solve_r;
  hipError_t (*launch)(const KArgs& a, int mode, int prec, int block, int grid, hipStream_t st);
  hipError_t (*info)(int mode, int prec, int block, KernelInfo* out);
  void (*param_defaults)(double* p);
};

std::vector<const NetEntry*>& registry();

struct Registrar {
  explicit Registrar(const NetEntry* e) { registry().push_back(e); }
};

__global__ void k_setq(unsigned int* q, unsigned int v);

inline namespace ODV_VARIANT_NS {

template <class K>
inline hipError_t kinfo(K kern, int block, KernelInfo* out) {
  hipFuncAttributes at;
  hipError_t e = hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kern));
  if (e != hipSuccess) return e;
  int nb = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kern), block, 0);
  if (e != hipSuccess) return e;
  out->regs = at.numRegs;
  out->sgprs = 0;
  out->spill_bytes = (int)at.localSizeBytes;
  out->blocks_per_cu = nb;
  out->block = block;
  return hipSuccess;
}

template <class Net, class M, bool RCP>
inline hipError_t launch_t(const KArgs& a, int mode, int block, int grid, hipStream_t st) {
  if (a.ncells <= 0) return hipSuccess;
  if (mode == MODE_PERCELL) {
    const int b = block > 0 ? block : 256;
    const long long g = (a.ncells + b - 1) / b;
    hipLaunchKernelGGL((odv_k_percell<Net, M, RCP>), dim3((unsigned)g), dim3(b), 0, st, a);
  } else if (mode == MODE_PERSISTENT) {
    const int b = block > 0 ? block : 256;
    int g = grid;
    if (g <= 0) {
      int dev = 0, ncu = 0, nb = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void*>(odv_k_persistent<Net, M, RCP>), b, 0);
      if (e != hipSuccess) return e;
      g = ncu * (nb > 0 ? nb : 1);
    }
    long long need = (a.ncells + b - 1) / b;
    if (g > need) g = (int)need;
    hipLaunchKernelGGL(k_setq, dim3(1), dim3(1), 0, st, a.queue, (unsigned)((long long)g * b));
    hipLaunchKernelGGL((odv_k_persistent<Net, M, RCP>), dim3(g), dim3(b), 0, st, a);
  } else {
    const int grp = a.group;
    const int b = ((grp + 63) / 64) * 64;
    const long long g = (a.ncells + grp - 1) / grp;
    hipLaunchKernelGGL((odv_k_lockstep<Net, M, RCP>), dim3((unsigned)g), dim3(b), 0, st, a);
  }
  return hipGetLastError();
}

template <class Net>
hipError_t launch_net(const KArgs& a, int mode, int prec, int block, int grid, hipStream_t st) {
  if (prec == PREC_MIXED) return launch_t<Net, float, true>(a, mode, block, grid, st);
  return launch_t<Net, double, false>(a, mode, block, grid, st);
}

template <class Net, class M, bool RCP>
inline hipError_t info_t(int mode, int block, KernelInfo* out) {
  if (mode == MODE_PERCELL) return kinfo(odv_k_percell<Net, M, RCP>, block > 0 ? block : 256, out);
  if (mode == MODE_PERSISTENT) return kinfo(odv_k_persistent<Net, M, RCP>, block > 0 ? block : 256, out);
  return kinfo(odv_k_lockstep<Net, M, RCP>, block > 0 ? block : 64, out);
}

template <class Net>
hipError_t info_net(int mode, int prec, int block, KernelInfo* out) {
  if (prec == PREC_MIXED) return info_t<Net, float, true>(mode, block, out);
  return info_t<Net, double, false>(mode, block, out);
}

}  // namespace ODV_VARIANT_NS
}  // namespace odv

#define ODV_REGISTER(NET)                                                                       \
  static const odv::NetEntry odv_entry_##NET = {  /* NOLINT */                                                \
      NET::ID, ODV_REPRO_FLAG, NET::NEQ, NET::NNZ, NET::NRATE, NET::NPARAM, NET::RATE_KIND, NET::TGAS,          \
      NET::FLOPS_RHS, NET::FLOPS_JAC, NET::FLOPS_LU, NET::FLOPS_SOLVE, NET::FLOPS_LU_R,         \
      NET::FLOPS_SOLVE_R, &odv::launch_net<NET>, &odv::info_net<NET>, &NET::param_defaults};                          \
  static odv::Registrar odv_reg_##NET(&odv_entry_##NET);

// Generated by odevec_amd.pre_odevec -- do not edit.
// network OREGO, ordering None, packaging CSC
// neq 3 (integrated 3), nnz(J) 7, nnz(P) 7, nnz(LU) 8 (full system 8), terms 9
// flops: rhs 17, jac 25, lu 8, solve 13, lu_r 11, solve_r 13
#pragma once
#include "odevec_netcommon.h"

struct odv_net_OREGO_NONE_CSC {
  static constexpr const char* ID = "OREGO/NONE/CSC";
  static constexpr int NEQ = 3;   // species in the caller's state
  static constexpr int NACT = 3;  // species integrated; the others stay constant
  // ACT[a] = species index of integrated species a
  static constexpr int ACT[3] = {0, 1, 2};
  // linear invariants over the integrated species: sum_a INV[k][a] f_a(y) = 0 for all y
  static constexpr int NINV = 0;
  static constexpr signed char INV[1][3] = {{}};
  ODV_HD static constexpr double inv(int k, int a) { return INV[k][a]; }
  static constexpr int NNZ = 8;
  static constexpr int NRATE = 1;
  static constexpr int NPARAM = 0;
  static constexpr int RATE_KIND = 0;
  static constexpr int TGAS = -1;
  static constexpr bool CROUT = true;
  // LUmethod 3: P by forward differences of the RHS (CalcJac_dense); AIDX[r*NACT + c]
  // = storage index of P_order(r, c), -1 outside the LU pattern (there the difference
  // is exactly zero: no RHS term of row r reads species c)
  static constexpr bool FD_JAC = false;
  // LUmethod 2: jac() writes the closed-form L and U factors of P_order (GetLU),
  // lu() does nothing and lu_r() only inverts the diagonal
  static constexpr int LU_METHOD = 1;
  static constexpr int FLOPS_RHS = 17, FLOPS_JAC = 25, FLOPS_LU = 8, FLOPS_SOLVE = 13;
  static constexpr int FLOPS_LU_R = 11, FLOPS_SOLVE_R = 13;
  // permutation Perm (0-based): P_order[i][j] = P[Perm[i]][Perm[j]]
  // Perm = {0, 1, 2}

  static inline void param_defaults(double* p) {
    (void)p;
  }

  ODV_HD static inline void rates(const double* __restrict__ y0, const double* __restrict__ prm,
                                 double* __restrict__ k) {
    (void)y0; (void)prm; (void)k;
  }
  static constexpr bool COOP = true;   // cooperative kernels available

  // cooperative kernels (species on lanes): tables in the oracle's order
  static constexpr int CO_MAXT = 4;   // most RHS terms of one species
  static constexpr int CO_MAXJ = 5;   // most Jacobian contributions of one row
  static constexpr short RT_OFF[4] = {
      0, 4, 7, 9,
  };
  static constexpr short RT_RATE[9] = {
      -1, -1, -1, -1, -1, -1, -1, -1, -1,
  };
  static constexpr double RT_COEF[9] = {
      77.27, -0.00064713625, -77.27, 77.27, -0.012941633234114146, -0.012941633234114146, 0.012941633234114146, 0.161,
      -0.161,
  };
  static constexpr unsigned char RT_NF[9] = {
      1, 2, 2, 1, 2, 1, 1, 1, 1,
  };
  static constexpr unsigned char RT_F[36] = {
      0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0,
      2, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0,
  };
  static constexpr short JT_OFF[10] = {
      0, 3, 5, 5, 6, 8, 9, 10, 10, 11,
  };
  static constexpr short JT_RATE[11] = {
      -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
  };
  static constexpr double JT_CE[11] = {
      77.27, -0.0012942725, -77.27, -77.27, 77.27, -0.012941633234114146, -0.012941633234114146, -0.012941633234114146,
      0.012941633234114146, 0.161, -0.161,
  };
  static constexpr unsigned char JT_NR[11] = {
      0, 1, 1, 1, 0, 1, 1, 0, 0, 0, 0,
  };
  static constexpr unsigned char JT_R[33] = {
      0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0,
      0, 0, 0, 0, 0, 0, 0, 0, 0,
  };
  static constexpr unsigned char PERM[3] = {
      0, 1, 2,
  };
  static constexpr unsigned char IPERM[3] = {
      0, 1, 2,
  };
  static constexpr unsigned long long LU_MASK[3] = {
      0x3ull, 0x7ull, 0x7ull,
  };
  static constexpr unsigned char RL_N[3] = {
      4, 3, 2,
  };
  static constexpr unsigned RL_W[12] = {
      131071u, 196607u, 131071u, 196607u, 131071u, 131071u, 196607u, 131071u, 0u, 131071u, 0u, 0u,
  };
  static constexpr unsigned RL_F[12] = {
      0u, 256u, 0u, 0u, 1u, 2u, 256u, 2u, 0u, 1u, 0u, 0u,
  };
  static constexpr double RL_C[12] = {
      77.27, -0.012941633234114146, 0.161, -0.00064713625, -0.012941633234114146, -0.161, -77.27, 0.012941633234114146,
      0.0, 77.27, 0.0, 0.0,
  };
  static constexpr unsigned char JL_N[3] = {
      5, 4, 2,
  };
  static constexpr unsigned JL_W[15] = {
      65535u, 131071u, 65535u, 131071u, 16908287u, 33619967u, 131071u, 16842751u, 0u, 16908287u, 33619967u, 0u,
      16842751u, 0u, 0u,
  };
  static constexpr unsigned JL_F[15] = {
      0u, 1u, 0u, 0u, 0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u,
      0u, 0u, 0u,
  };
  static constexpr double JL_C[15] = {
      77.27, -0.012941633234114146, 0.161, -0.0012942725, -0.012941633234114146, -0.161, -77.27, -0.012941633234114146,
      0.0, -77.27, 0.012941633234114146, 0.0, 77.27, 0.0, 0.0,
  };
  static constexpr bool LANE = true;   // per-lane kernels available
  // the reference's GetJac layout from this storage (see odevec_get_jac)
  static constexpr int JMAP_N = 8;
  static constexpr short JMAP[8] = {
      0, 1, 2, 3, 4, 5, 6, 7,
  };

  // k[r * S]: rate r of this cell (S = stride between a cell's rates)
  template <int S = 1>
  ODV_HD static inline void rhs(const double* __restrict__ y, const double* __restrict__ k,
                               double* __restrict__ f) {
#pragma clang fp contract(off)
    (void)k;
    {
      const double F = y[0];
      f[0] = (77.27 * F);
      f[2] = (0.161 * F);
    }
    {
      const double F = y[0] * y[0];
      f[0] = f[0] - 0.00064713625 * F;
    }
    {
      const double F = y[0] * y[1];
      f[0] = f[0] - 77.27 * F;
      f[1] = (-0.012941633234114146 * F);
    }
    {
      const double F = y[1];
      f[0] = f[0] + (77.27 * F);
      f[1] = f[1] - 0.012941633234114146 * F;
    }
    {
      const double F = y[2];
      f[1] = f[1] + (0.012941633234114146 * F);
      f[2] = f[2] - 0.161 * F;
    }
  }

  template <class M, int S = 1>
  ODV_HD static inline void jac(const double* __restrict__ y, const double* __restrict__ k,
                               double hb, M* __restrict__ a) {
    (void)k; (void)y; (void)hb;
    const double J0_0 = (77.27 * 1.0) - 0.0012942725 * y[0] - 77.27 * y[1];
    a[0] = (M)(1.0 - hb * J0_0);
    const double J1_0 = (-0.012941633234114146 * y[1]);
    a[1] = (M)(-(hb * J1_0));
    const double J2_0 = (0.161 * 1.0);
    a[2] = (M)(-(hb * J2_0));
    const double J0_1 = (-77.27 * y[0]) + (77.27 * 1.0);
    a[3] = (M)(-(hb * J0_1));
    const double J1_1 = (-0.012941633234114146 * y[0]) - 0.012941633234114146 * 1.0;
    a[4] = (M)(1.0 - hb * J1_1);
    a[5] = (M)(0.0);
    const double J1_2 = (0.012941633234114146 * 1.0);
    a[6] = (M)(-(hb * J1_2));
    const double J2_2 = (-0.161 * 1.0);
    a[7] = (M)(1.0 - hb * J2_2);
  }

  template <class M>
  ODV_HD static inline void lu(M* __restrict__ a) {
    a[3] = a[3] / a[0];
    a[4] = a[4] - a[3] * a[1];
    a[5] = a[5] - a[3] * a[2];
    a[6] = a[6] / a[4];
    a[7] = a[7] - a[6] * a[5];
  }

  template <class M>
  ODV_HD static inline void lu_r(M* __restrict__ a) {
    a[0] = (M)1 / a[0];
    a[3] = a[3] * a[0];
    a[4] = a[4] - a[3] * a[1];
    a[5] = a[5] - a[3] * a[2];
    a[4] = (M)1 / a[4];
    a[6] = a[6] * a[4];
    a[7] = a[7] - a[6] * a[5];
    a[7] = (M)1 / a[7];
  }

  template <class M, class V>
  ODV_HD static inline void solve(const M* __restrict__ a, V* __restrict__ x) {
    x[0] = x[0] / a[0];
    x[1] = x[1] - x[0] * a[1];
    x[2] = x[2] - x[0] * a[2];
    x[1] = x[1] / a[4];
    x[2] = x[2] - x[1] * a[5];
    x[2] = x[2] / a[7];
    x[1] = x[1] - x[2] * a[6];
    x[0] = x[0] - x[1] * a[3];
  }

  template <class M, class V>
  ODV_HD static inline void solve_r(const M* __restrict__ a, V* __restrict__ x) {
    x[0] = x[0] * a[0];
    x[1] = x[1] - x[0] * a[1];
    x[2] = x[2] - x[0] * a[2];
    x[1] = x[1] * a[4];
    x[2] = x[2] - x[1] * a[5];
    x[2] = x[2] * a[7];
    x[1] = x[1] - x[2] * a[6];
    x[0] = x[0] - x[1] * a[3];
  }
};

// Generated by odevec_amd.pre_odevec -- do not edit.
// network ROBER, ordering None, packaging CSC
// neq 3 (integrated 3), nnz(J) 7, nnz(P) 8, nnz(LU) 8 (full system 8), terms 6
// flops: rhs 8, jac 18, lu 9, solve 13, lu_r 12, solve_r 13
#pragma once
#include "odevec_netcommon.h"

struct odv_net_ROBER_NONE_CSC {
  static constexpr const char* ID = "ROBER/NONE/CSC";
  static constexpr int NEQ = 3;   // species in the caller's state
  static constexpr int NACT = 3;  // species integrated; the others stay constant
  // ACT[a] = species index of integrated species a
  static constexpr int ACT[3] = {0, 1, 2};
  // linear invariants over the integrated species: sum_a INV[k][a] f_a(y) = 0 for all y
  static constexpr int NINV = 1;
  static constexpr signed char INV[1][3] = {{1, 1, 1}};
  ODV_HD static constexpr double inv(int k, int a) { return INV[k][a]; }
  static constexpr int NNZ = 8;
  static constexpr int NRATE = 3;
  static constexpr int NPARAM = 3;
  static constexpr int RATE_KIND = 1;
  static constexpr int TGAS = -1;
  static constexpr bool CROUT = true;
  // LUmethod 3: P by forward differences of the RHS (CalcJac_dense); AIDX[r*NACT + c]
  // = storage index of P_order(r, c), -1 outside the LU pattern (there the difference
  // is exactly zero: no RHS term of row r reads species c)
  static constexpr bool FD_JAC = false;
  // LUmethod 2: jac() writes the closed-form L and U factors of P_order (GetLU),
  // lu() does nothing and lu_r() only inverts the diagonal
  static constexpr int LU_METHOD = 1;
  static constexpr int FLOPS_RHS = 8, FLOPS_JAC = 18, FLOPS_LU = 9, FLOPS_SOLVE = 13;
  static constexpr int FLOPS_LU_R = 12, FLOPS_SOLVE_R = 13;
  // permutation Perm (0-based): P_order[i][j] = P[Perm[i]][Perm[j]]
  // Perm = {0, 1, 2}

  static inline void param_defaults(double* p) {
    p[0] = 0.04;
    p[1] = 30000000.0;
    p[2] = 10000.0;
  }

  ODV_HD static inline void rates(const double* __restrict__ y0, const double* __restrict__ prm,
                                 double* __restrict__ k) {
    (void)y0;
    k[0] = prm[0];
    k[1] = prm[1];
    k[2] = prm[2];
  }
  static constexpr bool COOP = true;   // cooperative kernels available

  // cooperative kernels (species on lanes): tables in the oracle's order
  static constexpr int CO_MAXT = 3;   // most RHS terms of one species
  static constexpr int CO_MAXJ = 4;   // most Jacobian contributions of one row
  static constexpr short RT_OFF[4] = {
      0, 2, 5, 6,
  };
  static constexpr short RT_RATE[6] = {
      0, 2, 0, 1, 2, 1,
  };
  static constexpr double RT_COEF[6] = {
      -1.0, 1.0, 1.0, -1.0, -1.0, 1.0,
  };
  static constexpr unsigned char RT_NF[6] = {
      1, 2, 1, 2, 2, 2,
  };
  static constexpr unsigned char RT_F[24] = {
      0, 0, 0, 0, 1, 2, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 1, 2, 0, 0, 1, 1, 0, 0,
  };
  static constexpr short JT_OFF[10] = {
      0, 1, 2, 3, 4, 6, 7, 7, 8, 8,
  };
  static constexpr short JT_RATE[8] = {
      0, 2, 2, 0, 1, 2, 2, 1,
  };
  static constexpr double JT_CE[8] = {
      -1.0, 1.0, 1.0, 1.0, -2.0, -1.0, -1.0, 2.0,
  };
  static constexpr unsigned char JT_NR[8] = {
      0, 1, 1, 0, 1, 1, 1, 1,
  };
  static constexpr unsigned char JT_R[24] = {
      0, 0, 0, 2, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 2, 0, 0, 1, 0, 0, 1, 0, 0,
  };
  static constexpr unsigned char PERM[3] = {
      0, 1, 2,
  };
  static constexpr unsigned char IPERM[3] = {
      0, 1, 2,
  };
  static constexpr unsigned long long LU_MASK[3] = {
      0x7ull, 0x7ull, 0x6ull,
  };
  static constexpr unsigned char RL_N[3] = {
      2, 3, 1,
  };
  static constexpr unsigned RL_W[9] = {
      65536u, 65536u, 131073u, 131074u, 131073u, 0u, 0u, 131074u, 0u,
  };
  static constexpr unsigned RL_F[9] = {
      0u, 0u, 257u, 513u, 257u, 0u, 0u, 513u, 0u,
  };
  static constexpr double RL_C[9] = {
      -1.0, 1.0, 1.0, 1.0, -1.0, 0.0, 0.0, -1.0,
      0.0,
  };
  static constexpr unsigned char JL_N[3] = {
      3, 4, 1,
  };
  static constexpr unsigned JL_W[12] = {
      0u, 0u, 16842753u, 16842754u, 16842753u, 0u, 33619970u, 16842754u, 0u, 0u, 33619970u, 0u,
  };
  static constexpr unsigned JL_F[12] = {
      0u, 0u, 1u, 2u, 1u, 0u, 1u, 2u, 0u, 0u, 1u, 0u,
  };
  static constexpr double JL_C[12] = {
      -1.0, 1.0, 2.0, 1.0, -2.0, 0.0, 1.0, -1.0,
      0.0, 0.0, -1.0, 0.0,
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
      const double F = k[0 * S] * y[0];
      f[0] = (-F);
      f[1] = F;
    }
    {
      const double F = k[1 * S] * y[1] * y[1];
      f[1] = f[1] - F;
      f[2] = F;
    }
    {
      const double F = k[2 * S] * y[1] * y[2];
      f[0] = f[0] + F;
      f[1] = f[1] - F;
    }
  }

  template <class M, int S = 1>
  ODV_HD static inline void jac(const double* __restrict__ y, const double* __restrict__ k,
                               double hb, M* __restrict__ a) {
    (void)k; (void)y; (void)hb;
    const double J0_0 = (-(k[0 * S]));
    a[0] = (M)(1.0 - hb * J0_0);
    const double J1_0 = (k[0 * S]);
    a[1] = (M)(-(hb * J1_0));
    const double J0_1 = (k[2 * S] * y[2]);
    a[2] = (M)(-(hb * J0_1));
    const double J1_1 = (-2.0 * (k[1 * S] * y[1])) - (k[2 * S] * y[2]);
    a[3] = (M)(1.0 - hb * J1_1);
    const double J2_1 = (2.0 * (k[1 * S] * y[1]));
    a[4] = (M)(-(hb * J2_1));
    const double J0_2 = (k[2 * S] * y[1]);
    a[5] = (M)(-(hb * J0_2));
    const double J1_2 = (-(k[2 * S] * y[1]));
    a[6] = (M)(-(hb * J1_2));
    a[7] = (M)(1.0);
  }

  template <class M>
  ODV_HD static inline void lu(M* __restrict__ a) {
    a[2] = a[2] / a[0];
    a[3] = a[3] - a[2] * a[1];
    a[5] = a[5] / a[0];
    a[6] = a[6] - a[5] * a[1];
    a[6] = a[6] / a[3];
    a[7] = a[7] - a[6] * a[4];
  }

  template <class M>
  ODV_HD static inline void lu_r(M* __restrict__ a) {
    a[0] = (M)1 / a[0];
    a[2] = a[2] * a[0];
    a[3] = a[3] - a[2] * a[1];
    a[3] = (M)1 / a[3];
    a[5] = a[5] * a[0];
    a[6] = a[6] - a[5] * a[1];
    a[6] = a[6] * a[3];
    a[7] = a[7] - a[6] * a[4];
    a[7] = (M)1 / a[7];
  }

  template <class M, class V>
  ODV_HD static inline void solve(const M* __restrict__ a, V* __restrict__ x) {
    x[0] = x[0] / a[0];
    x[1] = x[1] - x[0] * a[1];
    x[1] = x[1] / a[3];
    x[2] = x[2] - x[1] * a[4];
    x[2] = x[2] / a[7];
    x[0] = x[0] - x[2] * a[5];
    x[1] = x[1] - x[2] * a[6];
    x[0] = x[0] - x[1] * a[2];
  }

  template <class M, class V>
  ODV_HD static inline void solve_r(const M* __restrict__ a, V* __restrict__ x) {
    x[0] = x[0] * a[0];
    x[1] = x[1] - x[0] * a[1];
    x[1] = x[1] * a[3];
    x[2] = x[2] - x[1] * a[4];
    x[2] = x[2] * a[7];
    x[0] = x[0] - x[2] * a[5];
    x[1] = x[1] - x[2] * a[6];
    x[0] = x[0] - x[1] * a[2];
  }
};

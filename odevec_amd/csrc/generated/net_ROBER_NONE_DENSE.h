// Generated by odevec_amd.pre_odevec -- do not edit.
// network ROBER, ordering None, packaging DENSE
// neq 3 (integrated 3), nnz(J) 7, nnz(P) 8, nnz(LU) 8 (full system 8), terms 6
// flops: rhs 8, jac 18, lu 8, solve 13, lu_r 11, solve_r 13
#pragma once
#include "odevec_netcommon.h"

struct odv_net_ROBER_NONE_DENSE {
  static constexpr const char* ID = "ROBER/NONE/DENSE";
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
  static constexpr bool CROUT = false;
  // LUmethod 3: P by forward differences of the RHS (CalcJac_dense); AIDX[r*NACT + c]
  // = storage index of P_order(r, c), -1 outside the LU pattern (there the difference
  // is exactly zero: no RHS term of row r reads species c)
  static constexpr bool FD_JAC = false;
  // LUmethod 2: jac() writes the closed-form L and U factors of P_order (GetLU),
  // lu() does nothing and lu_r() only inverts the diagonal
  static constexpr int LU_METHOD = 1;
  static constexpr int FLOPS_RHS = 8, FLOPS_JAC = 18, FLOPS_LU = 8, FLOPS_SOLVE = 13;
  static constexpr int FLOPS_LU_R = 11, FLOPS_SOLVE_R = 13;
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
  static constexpr bool COOP = false;
  static constexpr bool LANE = true;   // per-lane kernels available
  // the reference's GetJac layout from this storage (see odevec_get_jac)
  static constexpr int JMAP_N = 9;
  static constexpr short JMAP[9] = {
      0, 1, -1, 2, 3, 4, 5, 6, 7,
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
    a[1] = a[1] / a[0];
    a[3] = a[3] - a[1] * a[2];
    a[6] = a[6] - a[1] * a[5];
    a[4] = a[4] / a[3];
    a[7] = a[7] - a[4] * a[6];
  }

  template <class M>
  ODV_HD static inline void lu_r(M* __restrict__ a) {
    a[0] = (M)1 / a[0];
    a[1] = a[1] * a[0];
    a[3] = a[3] - a[1] * a[2];
    a[6] = a[6] - a[1] * a[5];
    a[3] = (M)1 / a[3];
    a[4] = a[4] * a[3];
    a[7] = a[7] - a[4] * a[6];
    a[7] = (M)1 / a[7];
  }

  template <class M, class V>
  ODV_HD static inline void solve(const M* __restrict__ a, V* __restrict__ x) {
    x[1] = x[1] - (a[1] * x[0]);
    x[2] = x[2] - (a[4] * x[1]);
    x[2] = x[2] / a[7];
    x[1] = (x[1] - (a[6] * x[2])) / a[3];
    x[0] = (x[0] - (a[2] * x[1] + a[5] * x[2])) / a[0];
  }

  template <class M, class V>
  ODV_HD static inline void solve_r(const M* __restrict__ a, V* __restrict__ x) {
    x[1] = x[1] - (a[1] * x[0]);
    x[2] = x[2] - (a[4] * x[1]);
    x[2] = x[2] * a[7];
    x[1] = (x[1] - (a[6] * x[2])) * a[3];
    x[0] = (x[0] - (a[2] * x[1] + a[5] * x[2])) * a[0];
  }
};

// Generated by odevec_amd.pre_odevec -- do not edit.
// network ROBER, ordering None, packaging DENSE
// neq 3 (integrated 3), nnz(J) 7, nnz(P) 8, nnz(LU) 8 (full system 8), terms 6
// flops: rhs 8, jac 30, lu 0, solve 13, lu_r 3, solve_r 13
#pragma once
#include "odevec_netcommon.h"

struct odv_net_ROBER_NONE_DENSE_LU2 {
  static constexpr const char* ID = "ROBER/NONE/DENSE/LU2";
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
  static constexpr int LU_METHOD = 2;
  static constexpr int FLOPS_RHS = 8, FLOPS_JAC = 30, FLOPS_LU = 0, FLOPS_SOLVE = 13;
  static constexpr int FLOPS_LU_R = 3, FLOPS_SOLVE_R = 13;
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
    const double u3 = hb * k[0 * S];
    const double u4 = 1.0 + u3;
    const double u6 = hb * k[2 * S];
    const double u8 = u6 * y[2];
    const double u9 = -u8;
    const double u11 = u6 * y[1];
    const double u12 = -u11;
    const double u13 = u3 / u4;
    const double u14 = -u13;
    const double u15 = k[2 * S] * y[2];
    const double u16 = -u15;
    const double u18 = k[1 * S] * y[1];
    const double u20 = 2.0 * u18;
    const double u21 = u16 - u20;
    const double u22 = hb * u21;
    const double u23 = 1.0 - u22;
    const double u24 = k[0 * S] * k[2 * S];
    const double u25 = u24 * y[2];
    const double u26 = hb * hb;
    const double u27 = u25 * u26;
    const double u28 = u27 / u4;
    const double u29 = u23 - u28;
    const double u30 = u24 * y[1];
    const double u31 = u30 * u26;
    const double u32 = u31 / u4;
    const double u33 = u11 - u32;
    const double u34 = hb * k[1 * S];
    const double u35 = u34 * y[1];
    const double u36 = 2.0 * u35;
    const double u37 = u36 / u29;
    const double u38 = -u37;
    const double u39 = u35 * u33;
    const double u40 = 2.0 * u39;
    const double u41 = u40 / u29;
    const double u42 = 1.0 + u41;
    a[0] = (M)(u4);
    a[2] = (M)(u9);
    a[5] = (M)(u12);
    a[1] = (M)(u14);
    a[3] = (M)(u29);
    a[6] = (M)(u33);
    a[4] = (M)(u38);
    a[7] = (M)(u42);
  }

  template <class M>
  ODV_HD static inline void lu(M* __restrict__ a) {
    (void)a;
  }

  template <class M>
  ODV_HD static inline void lu_r(M* __restrict__ a) {
    a[0] = (M)1 / a[0];
    a[3] = (M)1 / a[3];
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

// Generated by odevec_amd.pre_odevec -- do not edit.
// network OREGO, ordering None, packaging DENSE
// neq 3 (integrated 3), nnz(J) 7, nnz(P) 7, nnz(LU) 8 (full system 8), terms 9
// flops: rhs 17, jac 35, lu 0, solve 13, lu_r 3, solve_r 13
#pragma once
#include "odevec_netcommon.h"

struct odv_net_OREGO_NONE_DENSE_LU2 {
  static constexpr const char* ID = "OREGO/NONE/DENSE/LU2";
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
  static constexpr bool CROUT = false;
  // LUmethod 3: P by forward differences of the RHS (CalcJac_dense); AIDX[r*NACT + c]
  // = storage index of P_order(r, c), -1 outside the LU pattern (there the difference
  // is exactly zero: no RHS term of row r reads species c)
  static constexpr bool FD_JAC = false;
  // LUmethod 2: jac() writes the closed-form L and U factors of P_order (GetLU),
  // lu() does nothing and lu_r() only inverts the diagonal
  static constexpr int LU_METHOD = 2;
  static constexpr int FLOPS_RHS = 17, FLOPS_JAC = 35, FLOPS_LU = 0, FLOPS_SOLVE = 13;
  static constexpr int FLOPS_LU_R = 3, FLOPS_SOLVE_R = 13;
  // permutation Perm (0-based): P_order[i][j] = P[Perm[i]][Perm[j]]
  // Perm = {0, 1, 2}

  static inline void param_defaults(double* p) {
    (void)p;
  }

  ODV_HD static inline void rates(const double* __restrict__ y0, const double* __restrict__ prm,
                                 double* __restrict__ k) {
    (void)y0; (void)prm; (void)k;
  }
  static constexpr bool COOP = false;
  static constexpr bool LANE = true;   // per-lane kernels available
  // the reference's GetJac layout from this storage (see odevec_get_jac)
  static constexpr int JMAP_N = 9;
  static constexpr short JMAP[9] = {
      0, 1, 2, 3, 4, 5, -1, 6, 7,
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
    const double u5 = 0.0012942725 * y[0];
    const double u6 = 77.27 - u5;
    const double u8 = 77.27 * y[1];
    const double u9 = u6 - u8;
    const double u10 = hb * u9;
    const double u11 = 1.0 - u10;
    const double u12 = 77.27 * y[0];
    const double u13 = 77.27 - u12;
    const double u14 = hb * u13;
    const double u15 = -u14;
    const double u16 = hb * y[1];
    const double u18 = 0.012941633234114146 * u16;
    const double u19 = u18 / u11;
    const double u20 = -0.012941633234114146;
    const double u21 = 0.012941633234114146 * y[0];
    const double u22 = u20 - u21;
    const double u23 = hb * u22;
    const double u24 = 1.0 - u23;
    const double u25 = hb * hb;
    const double u26 = y[1] * u25;
    const double u27 = u26 * u13;
    const double u28 = 0.012941633234114146 * u27;
    const double u29 = u28 / u11;
    const double u30 = u24 + u29;
    const double u31 = 0.012941633234114146 * hb;
    const double u32 = -u31;
    const double u34 = 0.161 * hb;
    const double u35 = u34 / u11;
    const double u36 = -u35;
    const double u37 = u25 * u13;
    const double u38 = 0.161 * u37;
    const double u39 = u11 * u30;
    const double u40 = u38 / u39;
    const double u41 = -u40;
    const double u42 = 1.0 + u34;
    const double u43 = u25 * hb;
    const double u44 = u43 * u13;
    const double u46 = 0.0020836029506923776 * u44;
    const double u47 = u46 / u39;
    const double u48 = u42 - u47;
    a[0] = (M)(u11);
    a[3] = (M)(u15);
    a[1] = (M)(u19);
    a[4] = (M)(u30);
    a[6] = (M)(u32);
    a[2] = (M)(u36);
    a[5] = (M)(u41);
    a[7] = (M)(u48);
  }

  template <class M>
  ODV_HD static inline void lu(M* __restrict__ a) {
    (void)a;
  }

  template <class M>
  ODV_HD static inline void lu_r(M* __restrict__ a) {
    a[0] = (M)1 / a[0];
    a[4] = (M)1 / a[4];
    a[7] = (M)1 / a[7];
  }

  template <class M, class V>
  ODV_HD static inline void solve(const M* __restrict__ a, V* __restrict__ x) {
    x[1] = x[1] - (a[1] * x[0]);
    x[2] = x[2] - (a[2] * x[0] + a[5] * x[1]);
    x[2] = x[2] / a[7];
    x[1] = (x[1] - (a[6] * x[2])) / a[4];
    x[0] = (x[0] - (a[3] * x[1])) / a[0];
  }

  template <class M, class V>
  ODV_HD static inline void solve_r(const M* __restrict__ a, V* __restrict__ x) {
    x[1] = x[1] - (a[1] * x[0]);
    x[2] = x[2] - (a[2] * x[0] + a[5] * x[1]);
    x[2] = x[2] * a[7];
    x[1] = (x[1] - (a[6] * x[2])) * a[4];
    x[0] = (x[0] - (a[3] * x[1])) * a[0];
  }
};

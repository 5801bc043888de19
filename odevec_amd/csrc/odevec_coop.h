// Cooperative kernels: one cell per group of G lanes, one species per lane.
//
// The per-lane kernels keep a whole cell in one lane's registers, which
// limits them to networks of ~16 species and to one wavefront per SIMD. Here
// a cell's species are spread over G lanes (G = the power of two >= its
// integrated species, <= 64; groups never straddle a wavefront):
//
//  * lane l holds species l's Nordsieck history, correction, weight and
//    previous correction, and row l of the permuted iteration matrix P and
//    its factors (dense row, structural zeros kept as exact zeros) -- in its
//    registers, or, for one-wavefront groups in the mixed form, in the
//    group's LDS block (CoopKind below);
//  * the cell's rates, full state, parameters, the active species' current
//    values and one row of U (the broadcast row of the factorization) sit in
//    the group's LDS block;
//  * the integrator is the same odv_attempt as the per-lane kernels (the
//    reference's SolveLinearSystem), instantiated for a one-species lane
//    (CoopNet adapter) with a group whose species_sum adds the lanes'
//    partial norms in species order and whose votes are over the group;
//  * the RHS and Jacobian are table-driven gathers (each lane evaluates the
//    terms of its own species / P row, in the oracle's order), the Crout
//    factorization is right-looking over rows with U rows broadcast through
//    LDS, and the triangular solves pass the pivot value by shuffle.
//
// Every floating-point operation a cell sees is the oracle's, in the same
// order (only exact zeros of the dense storage are skipped), so the
// reproducible build equals the oracle bit for bit here too.
#pragma once
#include <type_traits>
#include "odevec_kernels.h"

namespace odv {
inline namespace ODV_VARIANT_NS {

extern __shared__ double odv_coop_lds[];   // groups' LDS blocks (dynamic size)

constexpr int COOP_BLOCK = 256;   // lanes per workgroup of the cooperative kernel

template <int G>
__device__ inline int coop_l() { return (int)(threadIdx.x & (G - 1)); }
template <int G>
__device__ inline int coop_base() { return (int)((threadIdx.x & 63) & ~(G - 1)); }
template <int G>
__device__ inline int coop_grp() { return (int)(threadIdx.x / G); }

// orders the LDS accesses of one wavefront's lanes (a group lives in one
// wavefront, whose LDS operations complete in program order)
__device__ inline void odv_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// value of lane `base + j` of this lane's group: a scalar read of that lane
// when the group is the whole wavefront (j is then uniform), else a shuffle
template <int G>
__device__ inline double coop_bcast(double v, int base, int j) {
  if constexpr (G == 64) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffull), j);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  } else {
    return __shfl(v, base + j, 64);
  }
}

template <int N>
struct CoopWidth {
  static constexpr int G = N <= 1 ? 1 : N <= 2 ? 2 : N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
};

// a group's LDS block (doubles)
template <class Net>
struct CoopLayout {
  static constexpr int K = 0;                                         // rates [NRATE]
  static constexpr int YF = K + Net::NRATE;                           // the cell's full state [NEQ]
  static constexpr int PR = YF + Net::NEQ;                            // parameters
  static constexpr int YA = PR + (Net::NPARAM > 0 ? Net::NPARAM : 1);  // integrated species' values [NACT]
  static constexpr int UR = YA + Net::NACT;                           // broadcast row [NACT]
  static constexpr int WORDS = UR + Net::NACT;
};

// The generated network seen by a one-species lane. MR_: void, each lane's
// row of P and the factors in its registers (a[0..NB-1]); else their element
// type, the rows in the group's LDS block: entry (row l, column j) at
// rows()[j*RS + l] (column stride G + 1: a lane's own entries and a row read
// across the lanes both fall in distinct banks), over the U-row broadcast
// slot, which this form does not need.
template <class Net_, int G_, class MR_ = void>
struct CoopNet {
  using Base = Net_;
  using Y = CoopLayout<Net_>;
  static constexpr int G = G_;
  static constexpr int NB = Net_::NACT;    // species of the cell
  static constexpr const char* ID = Net_::ID;
  static constexpr int NEQ = Net_::NEQ;     // the norms divide by the caller's species count
  static constexpr int NACT = 1;            // species per lane
  static constexpr int ACT[1] = {0};
  static constexpr bool LROWS = !std::is_void<MR_>::value;
  using MR = typename std::conditional<LROWS, MR_, float>::type;
  static constexpr int RS = G + 1;
  // a group's LDS block (doubles)
  // (NB + 1 columns: column NB is a spare that the factorisation's batches
  // use for their empty slots)
  static constexpr int STRIDE = LROWS ? Y::UR + ((NB + 1) * RS * (int)sizeof(MR) + 7) / 8 : Y::WORDS;
  static constexpr int NNZ = LROWS ? 1 : NB;   // a lane's row of P / the factors in registers
  static constexpr bool FD_JAC = false;     // the symbolic Jacobian (tables below)
  static constexpr int NINV = Net_::NINV;   // the cell's linear invariants (mixed form's convergence test)
  static constexpr int NRATE = 1;           // the rates live in the group's LDS block
  static constexpr int NPARAM = Net_::NPARAM;
  static constexpr int RATE_KIND = Net_::RATE_KIND;
  static constexpr int TGAS = Net_::TGAS;
  static constexpr int FLOPS_RHS = Net_::FLOPS_RHS, FLOPS_JAC = Net_::FLOPS_JAC;
  static constexpr int FLOPS_LU = Net_::FLOPS_LU, FLOPS_SOLVE = Net_::FLOPS_SOLVE;
  static constexpr int FLOPS_LU_R = Net_::FLOPS_LU, FLOPS_SOLVE_R = Net_::FLOPS_SOLVE;
  static_assert(NB <= G && G <= 64, "a cell must fit one wavefront");

  __device__ static double* blk() { return odv_coop_lds + coop_grp<G>() * STRIDE; }
  __device__ static MR* rows() { return reinterpret_cast<MR*>(blk() + Y::UR); }

  // coefficient of this lane's species in invariant k (0 on padding lanes)
  __device__ static double inv(int k, int) {
    const int l = coop_l<G>();
    return l < NB ? (double)Base::INV[k][l] : 0.0;
  }

  // once per cell: lane 0 evaluates the rates into the LDS block
  __device__ static void rates(const double* y0, const double* prm, double* k) {
    k[0] = 0.0;
    if (coop_l<G>() == 0) Base::rates(y0, prm, blk() + Y::K);
    odv_wave_sync();
  }

  // publish the lanes' species values, then each lane gathers its species' terms
  template <int S = 1>
  __device__ __forceinline__ static void rhs(const double* y, const double* k, double* f) {
    // each term rounded before it is added, as in the per-lane RHS, so the
    // rows' sums see the same rounded fluxes
#pragma clang fp contract(off)
    (void)k;
    double* b = blk();
    const int l = coop_l<G>();
    if (l < NB) b[Y::YA + l] = y[0];
    odv_wave_sync();
    const double* kk = b + Y::K;
    const double* ya = b + Y::YA;
    // the lane-major tables: term q of every lane's species in consecutive words
    // (every lane runs CO_MAXT slots without branches; a slot past its
    // species' count reads the zero padding and is not selected)
    double acc = 0.0;
    const int row = l < NB ? l : 0;
    const int nt = l < NB ? (int)Base::RL_N[row] : 0;
    for (int q = 0; q < Base::CO_MAXT; ++q) {
      const unsigned w = Base::RL_W[q * NB + row];
      const unsigned fw = Base::RL_F[q * NB + row];
      const double ce = Base::RL_C[q * NB + row];
      const int r = (int)(short)(w & 0xffffu);
      const int nf = (int)(w >> 16);
      const double kr = kk[r >= 0 ? r : 0];
      const double y0 = ya[fw & 0xffu];
      double m = r >= 0 ? kr : (nf > 0 ? y0 : 1.0);
      const int p0 = (r < 0 && nf > 0) ? 1 : 0;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const double yp = ya[(fw >> (8 * p)) & 0xffu];
        const double mp = m * yp;
        m = (p >= p0 && p < nf) ? mp : m;
      }
      const double v = ce * m;
      const double sum = acc + v;
      acc = q < nt ? ((q == 0) ? v : sum) : acc;
    }
    f[0] = acc;
    odv_wave_sync();
  }

  // entry (row, jj) of P = I - hb*J (permuted), from its Jacobian contributions
  __device__ __forceinline__ static double pentry(int row, int jj, const double* kk, const double* ya, double hb) {
    double acc = 0.0;
    bool has = false;
    const int c0 = Base::JT_OFF[row * NB + jj], c1 = Base::JT_OFF[row * NB + jj + 1];
    for (int c = c0; c < c1; ++c) {
      const int r = Base::JT_RATE[c];
      const int nr = Base::JT_NR[c];
      const unsigned char* fr = Base::JT_R + 3 * c;
      double m;
      int q = 0;
      if (r >= 0) m = kk[r];
      else if (nr > 0) { m = ya[fr[0]]; q = 1; }
      else m = 1.0;
      for (; q < nr; ++q) m = m * ya[fr[q]];
      const double v = Base::JT_CE[c] * m;
      acc = has ? acc + v : v;
      has = true;
    }
    const bool diag = row == jj;
    return has ? (diag ? 1.0 - hb * acc : -(hb * acc)) : (diag ? 1.0 : 0.0);
  }
  // entry (row, jj) of P from its accumulated Jacobian contributions
  __device__ __forceinline__ static double pfinal(int row, int jj, double acc, double hb) {
    return row == jj ? 1.0 - hb * acc : -(hb * acc);
  }

  // row l of P
  template <class M, int S = 1>
  __device__ __forceinline__ static void jac(const double* y, const double* k, double hb, M* a) {
    (void)k;
    double* b = blk();
    const int l = coop_l<G>();
    if (l < NB) b[Y::YA + l] = y[0];
    odv_wave_sync();
    const double* kk = b + Y::K;
    const double* ya = b + Y::YA;
    const int row = l < NB ? l : 0;
    if constexpr (LROWS) {
      // the row's contributions in one pass (lane-major tables, in the
      // order of its entries, then term order); entries without any keep
      // 1 on the diagonal and 0 elsewhere
      MR* A = rows();
      for (int jj = 0; jj < NB; ++jj) A[jj * RS + l] = (MR)(jj == row ? 1.0 : 0.0);
      // every lane runs CO_MAXJ slots without branches: a slot past the
      // row's count reads the tables' zero padding and changes nothing (its
      // results are not selected; a store it does not need goes to the
      // spare column NB), so the slots' loads can issue ahead
      const int nc = Base::JL_N[row];
      double acc = 0.0;
      int cur = -1;
      for (int q = 0; q < Base::CO_MAXJ; ++q) {
        const bool valid = q < nc;
        const unsigned w = Base::JL_W[q * NB + row];
        const unsigned fw = Base::JL_F[q * NB + row];
        const double ce = Base::JL_C[q * NB + row];
        const int r = (int)(short)(w & 0xffffu);
        const int nr = (int)((w >> 16) & 0xffu);
        const int jj = (int)(w >> 24);
        const double kr = kk[r >= 0 ? r : 0];
        const double y0 = ya[fw & 0xffu];
        double m = r >= 0 ? kr : (nr > 0 ? y0 : 1.0);
        const int p0 = (r < 0 && nr > 0) ? 1 : 0;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const double yp = ya[(fw >> (8 * p)) & 0xffu];
          const double mp = m * yp;
          m = (p >= p0 && p < nr) ? mp : m;
        }
        const double v = ce * m;
        const bool fresh = valid && jj != cur;
        const int to = (fresh && cur >= 0) ? cur : NB;
        A[to * RS + l] = (MR)pfinal(row, cur, acc, hb);
        const double sum = acc + v;
        acc = fresh ? v : (valid ? sum : acc);
        cur = fresh ? jj : cur;
      }
      if (cur >= 0) A[cur * RS + l] = (MR)pfinal(row, cur, acc, hb);
    } else {
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) a[jj] = (M)pentry(row, jj, kk, ya, hb);
    }
    odv_wave_sync();
  }

  // Crout (unit U, L keeps the diagonal: src/odevec.F90:929-971), rows on
  // lanes: at step k, lane k publishes row k, lane j > k forms
  // U(k,j) = A(k,j) / L(k,k), lane k keeps row k of U, and every lane i > k
  // with L(i,k) in the pattern updates A(i,j) -= L(i,k) U(k,j) over row k's
  // pattern (per entry, in increasing k: the oracle's sequence).
  // (rows in LDS: the same steps, with row k read in place -- entry (k, j)
  // at rows()[j*RS + k] -- so no row is published; step k's updates run over
  // row k's pattern bits in batches of 8 entries, loads before stores)
  template <class M>
  __device__ __forceinline__ static void lu(M* a) {
    if constexpr (LROWS) {
      (void)a;
      lu_rows();
    } else {
      lu_regs(a);
    }
  }
  __device__ __forceinline__ static void lu_rows() {
    MR* A = rows();
    const int l = coop_l<G>();
    const unsigned long long mine = l < NB ? Base::LU_MASK[l] : 0ull;
    for (int k = 0; k < NB; ++k) {
      const unsigned long long rowk = Base::LU_MASK[k];
      const MR lkk = A[k * RS + k];
      if (l > k && l < NB && ((rowk >> l) & 1ull)) A[l * RS + k] = A[l * RS + k] / lkk;
      odv_wave_sync();
      if (l > k && ((mine >> k) & 1ull)) {
        const MR lik = A[k * RS + l];
        unsigned long long m = rowk & ~((2ull << k) - 1ull);   // row k's pattern right of k (k = 63: none)
        // batches of 8 entries without branches (an empty slot works on the
        // spare column NB), so a batch's loads issue together
        while (m) {
          int js[8];
          MR u[8], v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            js[q] = m ? (int)__builtin_ctzll(m) : NB;
            m = m ? (m & (m - 1ull)) : 0ull;
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            u[q] = A[js[q] * RS + k];
            v[q] = A[js[q] * RS + l];
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) A[js[q] * RS + l] = v[q] - lik * u[q];
        }
      }
      odv_wave_sync();
    }
  }
  template <class M>
  __device__ __forceinline__ static void lu_regs(M* a) {
    double* b = blk();
    const int l = coop_l<G>();
    const int base = coop_base<G>();
    const unsigned long long mine = l < NB ? Base::LU_MASK[l] : 0ull;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const unsigned long long rowk = Base::LU_MASK[k];
      if (l == k) {
#pragma unroll
        for (int j = k + 1; j < NB; ++j)
          if ((rowk >> j) & 1ull) b[Y::UR + j] = (double)a[j];
      }
      odv_wave_sync();
      const M lkk = (M)coop_bcast<G>((double)a[k], base, k);
      if (l > k && l < NB && ((rowk >> l) & 1ull)) b[Y::UR + l] = (double)((M)b[Y::UR + l] / lkk);
      odv_wave_sync();
      const bool lik = l > k && ((mine >> k) & 1ull);
#pragma unroll
      for (int j = k + 1; j < NB; ++j) {
        if ((rowk >> j) & 1ull) {
          const M ukj = (M)b[Y::UR + j];
          if (l == k) a[j] = ukj;
          else if (lik) a[j] = a[j] - a[k] * ukj;
        }
      }
      odv_wave_sync();
    }
  }
  template <class M>
  __device__ __forceinline__ static void lu_r(M* a) { lu<M>(a); }

  // x (species l's residual) -> the correction; permutation in and out by shuffle
  template <class M, class V>
  __device__ __forceinline__ static void solve(const M* a, V* x) {
    const int l = coop_l<G>();
    const int base = coop_base<G>();
    const unsigned long long mine = l < NB ? Base::LU_MASK[l] : 0ull;
    V xp = (V)__shfl((double)x[0], base + (l < NB ? (int)Base::PERM[l] : 0), 64);
    if constexpr (LROWS) {
      (void)a;
      const MR* A = rows();
      // loads outside the lanes' conditions (selects, no branches), so the
      // unrolled steps' loads can issue ahead of the dependent chain
#pragma unroll 8
      for (int kk = 0; kk < NB; ++kk) {
        const V dk = (V)A[kk * RS + kk];
        const V lk = (V)A[kk * RS + l];
        const V xd = xp / dk;
        xp = (l == kk) ? xd : xp;
        const V mult = (V)coop_bcast<G>((double)xp, base, kk);
        const V xu = xp - mult * lk;
        xp = (l > kk && ((mine >> kk) & 1ull)) ? xu : xp;
      }
#pragma unroll 8
      for (int kk = NB - 1; kk >= 0; --kk) {
        const V uk = (V)A[kk * RS + l];
        const V mult = (V)coop_bcast<G>((double)xp, base, kk);
        const V xu = xp - mult * uk;
        xp = (l < kk && ((mine >> kk) & 1ull)) ? xu : xp;
      }
      x[0] = (V)__shfl((double)xp, base + (l < NB ? (int)Base::IPERM[l] : 0), 64);
      return;
    }
#pragma unroll
    for (int kk = 0; kk < NB; ++kk) {
      if (l == kk) xp = xp / (V)a[kk];
      const V mult = (V)coop_bcast<G>((double)xp, base, kk);
      if (l > kk && ((mine >> kk) & 1ull)) xp = xp - mult * (V)a[kk];
    }
#pragma unroll
    for (int kk = NB - 1; kk >= 0; --kk) {
      const V mult = (V)coop_bcast<G>((double)xp, base, kk);
      if (l < kk && ((mine >> kk) & 1ull)) xp = xp - mult * (V)a[kk];
    }
    x[0] = (V)__shfl((double)xp, base + (l < NB ? (int)Base::IPERM[l] : 0), 64);
  }
  template <class M, class V>
  __device__ __forceinline__ static void solve_r(const M* a, V* x) { solve<M, V>(a, x); }
};

template <class B, int G, class MR>
struct SpeciesMap<CoopNet<B, G, MR>> {
  static constexpr int NACT_ALL = B::NACT;
  static ODV_HD int all(int a) { return B::ACT[a]; }
  __device__ static int lane(int) {
    const int l = coop_l<G>();
    return B::ACT[l < B::NACT ? l : 0];
  }
};

// Votes and sums over a group's species lanes (per-cell semantics).
template <int G, int NB>
struct CoopGroup {
  __device__ double max_active(double v, bool) const { return v; }
  __device__ bool any_of(bool b) const {
    const unsigned long long ball = __ballot(b && coop_l<G>() < NB);
    const unsigned long long gm = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << coop_base<G>());
    return (ball & gm) != 0ull;
  }
  __device__ bool any_valid(bool b) const { return any_of(b); }
  __device__ bool any_active(bool b) const { return any_of(b); }
  // the largest of the lanes' values
  __device__ double species_max(double v) const {
    const int base = coop_base<G>();
    double m = coop_bcast<G>(v, base, 0);
#pragma unroll
    for (int j = 1; j < NB; ++j) m = fmax(m, coop_bcast<G>(v, base, j));
    return m;
  }
  // the lanes' partial sums added in species order: s = v_0; s = s + v_1; ...
  __device__ double species_sum(double v) const {
    const int base = coop_base<G>();
    double s = coop_bcast<G>(v, base, 0);
#pragma unroll
    for (int j = 1; j < NB; ++j) s = s + coop_bcast<G>(v, base, j);
    return s;
  }
};

// Where a network's cooperative kernel keeps P: one-wavefront groups (more
// than 32 species) keep the rows in LDS (16.9 KB of FP32 rows per wavefront,
// 8 one-wavefront workgroups per CU; 33.8 KB of FP64 rows, 4 per CU), so a
// lane's registers hold no row, the Jacobian is one pass over a row's
// contributions and the factorisation and solves run as loops rather than
// unrolled code; narrower groups keep their short rows in registers.
// Measured, SYNTH64, 65,536 cells, ms per launch, rows in registers / in
// LDS: mixed 198 / 98, FP64 283 / 185 (same step and iteration counts; in
// registers the kernels spilled 1.3 KB per lane and ran 36k instructions of
// unrolled code; in LDS, the loops' slots are branch-free so that their
// loads issue together: with a branch per slot, 137 / 250).
template <class Net, class M>
struct CoopKind {
  static constexpr int G = CoopWidth<Net::NACT>::G;
  static constexpr bool LROWS = G == 64;
  using CN = CoopNet<Net, G, typename std::conditional<LROWS, M, void>::type>;
  static constexpr int BLOCK = LROWS ? 64 : COOP_BLOCK;   // lanes per workgroup
  static constexpr int LDS_BYTES = (BLOCK / G) * CN::STRIDE * 8;
};

template <class Net, class M, bool RCP>
using CoopLane = Lane<Cfg<typename CoopKind<Net, M>::CN, M, RCP, 0, false>>;

template <class Net, class M>
constexpr int coop_lds_bytes() {
  return CoopKind<Net, M>::LDS_BYTES;
}

template <class Net, class M>
constexpr int coop_block() {
  return CoopKind<Net, M>::BLOCK;
}

// Warm restart of the cooperative kernel (no reference counterpart; the
// semantics of the per-lane kernels' Hist): cell i's record at
// hist[i*PLANES ...]. The group's scalars once; each lane's own state (the
// padding lanes' too, so a continued cell repeats their arithmetic exactly)
// in planes of G, lane l at [plane*G + l], so the lanes of a group write and
// read consecutive doubles.
template <class Net>
struct CoopHist {
  static constexpr int G = CoopWidth<Net::NACT>::G;
  static constexpr int NB = Net::NACT, NF = Net::NEQ, NP = Net::NPARAM;
  static constexpr int VALID = 0, TSTOP = 1, T = 2, DT = 3, RATE = 4, JCH = 5, CERR = 6;
  // q, steps since Jacobian, lju, failed tests, conv fails, flags, failed
  // invariant tests, convergence failures
  static constexpr int INTS = 7, NINT = 8;
  static constexpr int Y = INTS + NINT;        // the state returned (all NEQ species)
  static constexpr int P = Y + NF;             // the parameters
  static constexpr int Z = P + NP;             // Nordsieck column c of lane l: Z + c*G + l
  static constexpr int W = Z + NCOL * G;       // error weight
  static constexpr int EO = W + G;             // previous accepted correction
  static constexpr int FZ = EO + G;            // frozen-species flags (bit 3 is the lane's own)
  static constexpr int A = FZ + G;             // entry j of lane l's row of the factors: A + j*G + l
  static constexpr int K = A + NB * G;         // rate coefficients
  static constexpr int PLANES = K + Net::NRATE;
};

// any lane of this lane's group
template <int G>
__device__ inline bool coop_any(bool b) {
  const unsigned long long ball = __ballot(b);
  const unsigned long long gm = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << coop_base<G>());
  return (ball & gm) != 0ull;
}

// continue cell i from its history if the call starts where it stopped, on
// the state and parameters it returned (in the group's LDS block)
template <class Net, class L>
__device__ inline bool coop_hist_resume(const KArgs& a, long long i, L& s, const double* b, double t0,
                                        double ts) {
  using H = CoopHist<Net>;
  using Y = CoopLayout<Net>;
  using M = typename L::M;
  constexpr int G = H::G;
  const int l = coop_l<G>();
  const double* h = a.hist + i * H::PLANES;
  bool ok = h[H::VALID] == 1.0 && h[H::TSTOP] == t0;
  for (int j = l; j < H::NF; j += G) ok = ok && h[H::Y + j] == b[Y::YF + j];
  for (int p = l; p < H::NP; p += G) ok = ok && h[H::P + p] == b[Y::PR + p];
  if (coop_any<G>(!ok)) return false;
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
  s.cnt.invfail = (int)h[H::INTS + 6];
  s.cnt.convfail = (int)h[H::INTS + 7];
  s.status = ST_RUNNING;
  s.tstop = ts;
#pragma unroll
  for (int c = 0; c < NCOL; ++c) s.set_z(c, 0, h[H::Z + c * G + l]);
  s.set_w(0, h[H::W + l]);
  s.set_eo(0, h[H::EO + l]);
  s.en[0] = 0.0;
  s.fz = (int)h[H::FZ + l];
  if constexpr (L::Net::LROWS) {
    typename L::Net::MR* A = L::Net::rows();
    for (int j = 0; j < H::NB; ++j) A[j * L::Net::RS + l] = (typename L::Net::MR)h[H::A + j * G + l];
  } else {
#pragma unroll
    for (int j = 0; j < H::NB; ++j) s.a[j] = (M)h[H::A + j * G + l];
  }
  s.kl[0] = 0.0;   // the rates live in the group's LDS block
  double* bw = const_cast<double*>(b);
  for (int r = l; r < Net::NRATE; r += G) bw[Y::K + r] = h[H::K + r];
  odv_wave_sync();
  return true;
}

// after a successful solve: yv this lane's result (its species', if active)
template <class Net, class L>
__device__ inline void coop_hist_save(const KArgs& a, long long i, const L& s, double* b, double yv) {
  using H = CoopHist<Net>;
  using Y = CoopLayout<Net>;
  constexpr int G = H::G;
  const int l = coop_l<G>();
  double* h = a.hist + i * H::PLANES;
  if (s.status != ST_OK) {
    if (l == 0) h[H::VALID] = 0.0;
    return;
  }
  // the state returned: the block's full state with the integrated species' results
  if (l < H::NB) b[Y::YF + Net::ACT[l]] = yv;
  odv_wave_sync();
  for (int j = l; j < H::NF; j += G) h[H::Y + j] = b[Y::YF + j];
  for (int p = l; p < H::NP; p += G) h[H::P + p] = b[Y::PR + p];
  for (int r = l; r < Net::NRATE; r += G) h[H::K + r] = b[Y::K + r];
#pragma unroll
  for (int c = 0; c < NCOL; ++c) h[H::Z + c * G + l] = s.z(c, 0);
  h[H::W + l] = s.w(0);
  h[H::EO + l] = s.eo(0);
  h[H::FZ + l] = (double)s.fz;
  if constexpr (L::Net::LROWS) {
    const typename L::Net::MR* A = L::Net::rows();
    for (int j = 0; j < H::NB; ++j) h[H::A + j * G + l] = (double)A[j * L::Net::RS + l];
  } else {
#pragma unroll
    for (int j = 0; j < H::NB; ++j) h[H::A + j * G + l] = (double)s.a[j];
  }
  if (l == 0) {
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
    h[H::INTS + 5] = (double)((s.first ? 1 : 0) | (s.upd_jac ? 2 : 0) | (s.neglect_up ? 4 : 0));
    h[H::INTS + 6] = (double)s.cnt.invfail;
    h[H::INTS + 7] = (double)s.cnt.convfail;
    h[H::VALID] = 1.0;
  }
}

// Persistent: each group takes cells from a queue (its first cell is its
// global group id; the counter starts at the number of groups launched)
// until none is left, so a long cell does not hold a finished group idle.
// Groups of a workgroup are independent (no workgroup barrier) and every
// group-wide decision is a group vote, so a group's lanes follow one control
// path. Each group's loop ends: every cell is bounded by max_attempts.
// The network functions above are inlined (__forceinline__): called out of
// line, a lane's register row (passed by pointer) lives in scratch memory.
// Narrower groups (<= 32 species) keep the compiler's own register budget:
// their rows are short. One-wavefront groups keep the rows in LDS
// (CoopKind), which admits 2 (FP32 rows) or 1 (FP64) wavefronts per SIMD;
// they are compiled for a 256-register budget.
constexpr int COOP_WAVES_LROWS = 2;

template <class Net, class M>
constexpr int coop_waves() {
  return CoopKind<Net, M>::LROWS ? COOP_WAVES_LROWS : 1;
}

// HIST: the variant that continues cells from a warm-restart history and
// saves it (for networks that only this kernel runs)
template <class Net, class M, bool RCP, int HIST = 0>
__global__ void __launch_bounds__((coop_block<Net, M>()))
__attribute__((amdgpu_waves_per_eu(coop_waves<Net, M>())))
odv_k_coop(KArgs a) {
  constexpr int G = CoopWidth<Net::NACT>::G;
  using CN = typename CoopKind<Net, M>::CN;
  using L = CoopLane<Net, M, RCP>;
  using Y = CoopLayout<Net>;
  const int l = coop_l<G>();
  const int base = coop_base<G>();
  double* b = CN::blk();
  long long i = (long long)blockIdx.x * (blockDim.x / G) + coop_grp<G>();
  CoopGroup<G, Net::NACT> g;
  double fl64 = 0.0, fl32 = 0.0;
  while (i < a.ncells) {
    for (int j = l; j < Net::NEQ; j += G) b[Y::YF + j] = a.y[(long long)j * a.ncells + i];
    for (int p = l; p < Net::NPARAM; p += G) b[Y::PR + p] = a.prm[p * a.ps_plane + i * a.ps_cell];
    odv_wave_sync();
    if (a.times && l == 0) a.times[i] = (long long)wall_clock64();
    const double t0 = a.t0[i * a.t0_s];
    const double ts = a.tstop[i * a.ts_s];
    const double dh = a.dt0[i * a.dt_s];
    L s;
    s.active = l < Net::NACT;
    s.cnt = Counters{0, 0, 0, 0, 0, 0, 0, 0};
    s.dt = dh;
    if ((a.mask && !a.mask[i]) || !(t0 < ts)) {
      if (l == 0) {
        if (a.status) a.status[i] = ST_SKIPPED;
        if (a.dt_out) a.dt_out[i] = dh;
        if (a.counters) odv_store_counters(a, i, s.cnt);
      }
    } else {
      bool resumed = false;
      if constexpr (HIST) {
        if (a.hist) resumed = coop_hist_resume<Net>(a, i, s, b, t0, ts);
      }
      if (!resumed) odv_begin(s, b + Y::YF, b + Y::PR, t0, ts, dh, a.ctl, g);
      if (resumed && !(s.t < s.tstop)) {
        s.status = ST_OK;   // the last step already passed the new t_stop: interpolate
      } else {
        while (!odv_attempt(s, a.ctl, g)) {
        }
      }
      double yv[1];
      odv_result(s, yv);
      if (l < Net::NACT) a.y[(long long)Net::ACT[l] * a.ncells + i] = yv[0];
      if constexpr (HIST) {
        if (a.hist) coop_hist_save<Net>(a, i, s, b, yv[0]);
      }
      if (l == 0) {
        if (a.status) a.status[i] = s.status;
        if (a.dt_out) a.dt_out[i] = s.dt;
        if (a.counters) odv_store_counters(a, i, s.cnt);
        if (a.flops) {
          double f64, f32;
          odv_flops<Lane<Cfg<Net, M, RCP, 0, false>>>(s.cnt, f64, f32);
          fl64 += f64;
          fl32 += f32;
        }
      }
    }
    if (a.times && l == 0) a.times[a.ncells + i] = (long long)wall_clock64();
    odv_wave_sync();   // the group's LDS block is reloaded for the next cell
    unsigned long long nxt = 0;
    if (l == 0) nxt = atomicAdd(a.queue, 1ull);
    i = (long long)(unsigned long long)__shfl((long long)nxt, base, 64);
  }
  if (a.flops && l == 0 && (fl64 != 0.0 || fl32 != 0.0)) {
    atomicAdd(&a.flops[0], fl64);
    atomicAdd(&a.flops[1], fl32);
  }
}


}  // namespace ODV_VARIANT_NS
}  // namespace odv

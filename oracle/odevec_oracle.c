/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference OdeVec solver (src/odevec.F90) and
 * of its generated network plugin (src/odevec_commons.F90), used as the
 * checker of the HIP product. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg call it; the product never does.
 *
 * Structure follows the reference: one solve works on a vector of G cells
 * stored cell-fastest (y[j*G + i], like the Fortran y(nvector,neq)); every
 * norm is a max over the G cells (src/odevec.F90:669-702), so all cells
 * share one step size, order and verdict. G = 1 gives the per-cell
 * semantics. Networks are evaluated from the preprocessor's term table
 * (one monomial per term, canonical left-to-right order), the Jacobian by
 * differentiating those monomials (or, LUmethod 3, by CalcJac_dense's
 * forward differences of the RHS, src/odevec.F90:469-506), the factorizations with the CSC index
 * arrays exactly as the reference loops over them (src/odevec.F90:853-971)
 * or dense (src/odevec.F90:797-849, 898-922); for LUmethod 2 (GetLU) the
 * factors come from the preprocessor's closed-form program, executed
 * operation by operation (closed_lu below).
 *
 * Parity pins (see tests/): BDF l-vectors against their closed form,
 * LU factors (numeric and closed-form) against L*U = P, the closed-form
 * program against the exact rational factorization, RHS against the reference network
 * definitions' own values, rates against the reference formulas, whole
 * solves against an independent Radau integrator at rtol 1e-12, the
 * forward-difference Jacobian against the analytic one within its error
 * model (truncation + rounding).
 *
 * Build: make -C oracle (gcc, -O2 -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXORD 5
#define F32(v) ((double)(float)(v))

/* ------------------------------------------------------------ network */
typedef struct {
  int neq, nrate, nparam, rate_kind, tgas;
  int nterms;
  const int* t_row;      /* [nterms] */
  const double* t_coef;  /* [nterms] */
  const int* t_rate;     /* [nterms], -1 = none */
  const int* t_nfac;     /* [nterms] */
  const int* t_fac;      /* [nterms*4] factors, ascending with repeats */
  const int* perm;       /* [neq] P_order[i][j] = P[perm[i]][perm[j]] */
  int csc;               /* 1: CSC Crout, 0: dense Doolittle */
  int nnz;
  const int* row_index;  /* [nnz] */
  const int* u_col_start;/* [neq+1] */
  const int* l_col_start;/* [neq] */
  const double* arr_A;   /* arrhenius */
  const double* arr_Ea;
  int lu_method;         /* 1: GetJac (symbolic P); 2: GetLU (closed-form factors); 3: CalcJac_dense */
  /* LUmethod 2: the closed-form factors as a straight-line program over
   * y[a], k[a], hb, constants and earlier results (code 0 y, 1 k, 2 hb,
   * 3 constant, 4 add, 5 sub, 6 mul, 7 div, 8 negate), and its outputs
   * (row, col, node) in P_order coordinates */
  int lc_nops;
  const int* lc_code;
  const int* lc_a;
  const int* lc_b;
  const double* lc_c;
  int lc_nout;
  const int* lc_out;     /* [lc_nout*3] */
} onet_t;


/* Deterministic elementary functions: fixed sequences of correctly rounded
 * IEEE operations (the same published method as the product's: exp by
 * k*ln2 range reduction and a degree-13 Taylor polynomial, log by the atanh
 * series of s = (m-1)/(m+1), pow = exp(y log x)), so that the product's GPU
 * build can be compared with this oracle bit for bit. */
static double o_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  double kd = x * inv_ln2;
  kd = (kd >= 0.0) ? floor(kd + 0.5) : -floor(0.5 - kd);
  const int k = (int)kd;
  const double r = (x - kd * ln2_hi) - kd * ln2_lo;
  static const double c[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                               1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0,
                               1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
  double p = c[0];
  for (int i = 1; i < 14; ++i) p = p * r + c[i];
  return ldexp(p, k);
}

static double o_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int e = 0;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  for (int d = 21; d >= 3; d -= 2) p = p * s2 + 1.0 / (double)d;
  const double logm = 2.0 * s + (2.0 * s) * (s2 * p);
  const double ed = (double)e;
  return ed * 6.93147180369123816490e-01 + (logm + ed * 1.90821492927058770002e-10);
}

static double o_log10(double x) { return o_log(x) * 0.43429448190325182765; }

static double o_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (x == 0.0) return (y > 0.0) ? 0.0 : INFINITY;
  if (x < 0.0) return NAN;
  return o_exp(y * o_log(x));
}

double oracle_exp(double x) { return o_exp(x); }
double oracle_log(double x) { return o_log(x); }
double oracle_pow(double x, double y) { return o_pow(x, y); }

/* x**n by right-to-left binary exponentiation (same product order as the
 * product's rate code). */
static double ipow(double x, int n) {
  double r = 1.0;
  int have = 0;
  double b = x;
  while (n) {
    if (n & 1) { r = have ? r * b : b; have = 1; }
    n >>= 1;
    if (n) b = b * b;
  }
  return r;
}

static double lnpoly(const double* c, int n, double x) {
  double s = c[0];
  for (int i = 1; i < n; ++i) s = s + c[i] * ipow(x, i);
  return s;
}

/* KROME primordial rates, src/odevec_commons.F90:101-384, Tgas from the
 * state slot tgas; small from the state passed in (src/odevec_commons.F90:127-130). */
void oracle_primordial_rates(const double* y, int neq, int tgas, double* k) {
  double nmax = y[0];
  for (int j = 1; j < neq; ++j) nmax = fmax(nmax, y[j]);
  nmax = fmax(nmax, 1.0);
  const double small = 1e-40 / (nmax * nmax * nmax);
  const double T = y[tgas];
  const double Te = T * 8.617343e-5;
  const double lnTe = o_log(Te);
  const double T32 = T * F32(0.0033333333333333335);
  const double invT = 1.0 / T;
  const double invTe = 1.0 / Te;
  const double sq = sqrt(T);
  for (int i = 0; i < 38; ++i) k[i] = small;
  static const double c1[9] = {-32.71396786, 13.5365560, -5.73932875, 1.56315498, -0.28770560,
                               3.48255977e-2, -2.63197617e-3, 1.11954395e-4, -2.03914985e-6};
  static const double c3[10] = {-28.61303380689232, -0.7241125657826851, -0.02026044731984691,
                                -0.002380861877349834, -0.0003212605213188796, -0.00001421502914054107,
                                4.989108920299513e-6, 5.755614137575758e-7, -1.856767039775261e-8,
                                -3.071135243196595e-9};
  static const double c4[9] = {-44.09864886, 23.91596563, -10.7532302, 3.05803875, -0.56851189,
                               6.79539123e-2, -5.00905610e-3, 2.06723616e-4, -3.64916141e-6};
  static const double c7[9] = {-68.71040990212001, 43.93347632635, -18.48066993568, 4.701626486759002,
                               -0.7692466334492, 0.08113042097303, -0.005324020628287001,
                               0.0001975705312221, -3.165581065665e-6};
  static const double c11[8] = {-20.06913897587003, 0.2289800603272916, 0.03599837721023835,
                                -0.004555120027032095, -0.0003105115447124016, 0.0001073294010367247,
                                -8.36671960467864e-6, 2.238306228891639e-7};
  static const double c15[9] = {-24.24914687731536, 3.400824447095291, -3.898003964650152,
                                2.045587822403071, -0.5416182856220388, 0.0841077503763412,
                                -0.007879026154483455, 0.0004138398421504563, -9.36345888928611e-6};
  static const double c18[9] = {-18.01849334273, 2.360852208681, -0.2827443061704, 0.01623316639567,
                                -0.03365012031362999, 0.01178329782711, -0.001656194699504,
                                0.0001068275202678, -2.631285809207e-6};
  static const double c20[10] = {-20.37260896533324, 1.139449335841631, -0.1421013521554148,
                                 0.00846445538663, -0.0014327641212992, 0.0002012250284791,
                                 0.0000866396324309, -0.00002585009680264, 2.4555011970392e-6,
                                 -8.06838246118e-8};
  k[0] = small + o_exp(lnpoly(c1, 9, lnTe));
  if (T <= 5.5e3) k[1] = small + 3.92e-13 * o_pow(invTe, 0.6353);
  if (T > 5.5e3) k[2] = small + o_exp(lnpoly(c3, 10, lnTe));
  k[3] = small + o_exp(lnpoly(c4, 9, lnTe));
  if (T <= 9.28e3) k[4] = small + 3.92e-13 * o_pow(invTe, 0.6353);
  if (T > 9.28e3)
    k[5] = small + (1.54e-9 * (1.0 + 0.3 / o_exp(8.099328789667 * invTe)) /
                        (o_exp(40.49664394833662 * invTe) * o_pow(Te, 1.5)) +
                    3.92e-13 / o_pow(Te, 0.6353));
  k[6] = small + o_exp(lnpoly(c7, 9, lnTe));
  k[7] = small + (3.36e-10 / sq / o_pow(T / 1.e3, 0.2) / (1.0 + o_pow(T / 1.e6, 0.7)));
  k[8] = small + 6.77e-15 * o_pow(Te, 0.8779);
  if (T < 1160.0) k[9] = small + 1.43e-9;
  if (T > 1160.0) k[10] = small + o_exp(lnpoly(c11, 8, lnTe));
  if (T <= 6.7e3) k[11] = small + 1.85e-23 * o_pow(T, 1.8);
  if (T > 6.7e3) k[12] = small + 5.81e-16 * o_pow(T / 5.62e4, -0.6657 * o_log10(T / 5.62e4));
  k[13] = small + 6.0e-10;
  if (T > 3.48e3) k[14] = small + o_exp(lnpoly(c15, 9, lnTe));
  k[15] = small + 5.6e-11 * o_exp(-102124.0 * invT) * o_pow(T, 0.5);
  k[16] = small + 1.0670825e-10 * o_pow(Te, 2.012) * o_exp(-4.463 * invTe) / o_pow(1.0 + 0.2472 * Te, 3.512);
  k[17] = small + o_exp(lnpoly(c18, 9, lnTe));
  if (T <= 1.16e3) k[18] = small + 2.56e-9 * o_pow(Te, 1.78186);
  if (T > 1.16e3) k[19] = small + o_exp(lnpoly(c20, 10, lnTe));
  k[20] = small + 6.5e-9 / sqrt(Te);
  k[21] = small + 1.e-8 * o_pow(T, -0.4);
  if (T <= 6.17e2) k[22] = small + 1.e-8;
  if (T > 6.17e2) k[23] = small + 1.32e-6 * o_pow(T, -0.76);
  k[24] = small + 5.e-7 * sqrt(1.e2 * invT);
  if (T <= 3e2) k[25] = small + 1.3e-32 * o_pow(T32, -0.38);
  if (T > 3e2) k[26] = small + 1.3e-32 * o_pow(T32, -1.00);
  if (T <= 3e2) k[27] = small + 1.3e-32 * o_pow(T32, -0.38) / 8.0;
  if (T > 3e2) k[28] = small + 1.3e-32 * o_pow(T32, -1.00) / 8.0;
  if (T >= 5e1) k[29] = small + (2.00e-10 * o_pow(T, 0.402) * o_exp(-37.1 * invT) - 3.31e-17 * o_pow(T, 1.48));
  if (T >= 5e1)
    k[30] = small + (2.06e-10 * o_pow(T, F32(0.396)) * o_exp(-33.0 * invT) + 2.03e-9 * o_pow(T, -F32(0.332)));
  {
    const double l = o_log10(T);
    k[31] = small + 1.e-9 * (F32(0.417) + F32(0.846) * l - F32(0.137) * ipow(l, 2));
  }
  k[32] = small + 1.0e-9 * o_exp(-4.57e2 * invT);
  if (T <= 2.e3) {
    const double l = o_log10(T);
    k[33] = small + o_pow(10.0, -F32(56.4737) + F32(5.88886) * l + F32(7.19692) * ipow(l, 2) +
                                  F32(2.25069) * ipow(l, 3) - F32(2.16903) * ipow(l, 4) +
                                  F32(0.317887) * ipow(l, 5));
  }
  if (T > 2e3) k[34] = small + 3.17e-10 * o_exp(-5207.0 * invT);
  if (T > 2e2) k[35] = small + 5.25e-11 * o_exp(-4430.0 * invT + 1.739e5 * ipow(invT, 2));
  k[36] = small + 1.5e-9 * o_pow(T32, -0.1);
  k[37] = small + 3.6e-12 * o_pow(T / 300.0, -0.75);
}

/* rates of one cell: y0 its state, p its parameters */
static void cell_rates(const onet_t* n, const double* y0, const double* p, double* k) {
  switch (n->rate_kind) {
    case 1: for (int r = 0; r < n->nrate; ++r) k[r] = p[r]; break;
    case 2: oracle_primordial_rates(y0, n->neq, n->tgas, k); break;
    case 3: {
      const double invT = 1.0 / p[0];
      for (int r = 0; r < n->nrate; ++r) k[r] = n->arr_A[r] * o_exp(-(n->arr_Ea[r] * invT));
      break;
    }
    default: break;
  }
}

/* RHS of one cell from the term table (canonical order) */
void oracle_rhs(const onet_t* n, const double* y, const double* k, double* f) {
  int last = -1;
  for (int i = 0; i < n->neq; ++i) f[i] = 0.0;
  int* started = (int*)calloc((size_t)n->neq, sizeof(int));
  for (int t = 0; t < n->nterms; ++t) {
    const int* fac = n->t_fac + 4 * t;
    double m;
    int s = 0;
    if (n->t_rate[t] >= 0) m = k[n->t_rate[t]];
    else if (n->t_nfac[t] > 0) { m = y[fac[0]]; s = 1; }
    else m = 1.0;
    for (int q = s; q < n->t_nfac[t]; ++q) m = m * y[fac[q]];
    const double v = n->t_coef[t] * m;
    const int row = n->t_row[t];
    if (!started[row]) { f[row] = v; started[row] = 1; }
    else f[row] = f[row] + v;
    last = row;
  }
  (void)last;
  free(started);
}

/* dense P_order = I - hb*J (permuted), row-major [neq*neq], for one cell */
void oracle_p_dense(const onet_t* n, const double* y, const double* k, double hb, double* P) {
  const int N = n->neq;
  double* J = (double*)calloc((size_t)N * N, sizeof(double));
  char* has = (char*)calloc((size_t)N * N, 1);
  for (int t = 0; t < n->nterms; ++t) {
    const int* fac = n->t_fac + 4 * t;
    const int nf = n->t_nfac[t];
    for (int q = 0; q < nf; ++q) {
      const int j = fac[q];
      if (q > 0 && fac[q - 1] == j) continue;  /* distinct factors only */
      int e = 0;
      for (int r = 0; r < nf; ++r) e += fac[r] == j;
      /* reduced monomial: drop one occurrence of j */
      int red[4], nr = 0, dropped = 0;
      for (int r = 0; r < nf; ++r) {
        if (!dropped && fac[r] == j) { dropped = 1; continue; }
        red[nr++] = fac[r];
      }
      double m;
      int s = 0;
      if (n->t_rate[t] >= 0) m = k[n->t_rate[t]];
      else if (nr > 0) { m = y[red[0]]; s = 1; }
      else m = 1.0;
      for (int r = s; r < nr; ++r) m = m * y[red[r]];
      const double v = (n->t_coef[t] * e) * m;
      const int row = n->t_row[t];
      if (!has[row * N + j]) { J[row * N + j] = v; has[row * N + j] = 1; }
      else J[row * N + j] = J[row * N + j] + v;
    }
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const int R = n->perm[i], C = n->perm[j];
      double v;
      if (has[R * N + C]) v = (R == C) ? 1.0 - hb * J[R * N + C] : -(hb * J[R * N + C]);
      else v = (R == C) ? 1.0 : 0.0;
      P[i * N + j] = v;
    }
  free(J);
  free(has);
}

/* ------------------------------------------------------------ solver */
typedef struct {
  double rtol, atol, dt_min;
  int check_negatives, max_attempts;
} octl_t;

typedef struct {
  int attempts, steps, newton, jac, errfail, convfail, stepsel, qsum;
} ocnt_t;

/* one vector of G cells (the reference's type(odevec) instance) */
typedef struct {
  const onet_t* n;
  int G, N;
  double* yNS;   /* [(MAXORD+2)][N][G]: y_NS(i,j,k) at (k*N + j)*G + i */
  double *en, *enold, *den, *w, *rhs, *res, *LU, *k, *tmp;
  const char* mask;  /* [G] or NULL */
  double coeff[7][7]; /* coeff[q][k], q=1..6 */
  int order, swu, lju, fet, fc;
  int upd_jac, neglect_up, first_step;
  double rate, jchange, conv_err;
} ovec_t;

#define YNS(v, kk, j, i) (v)->yNS[((size_t)(kk) * (v)->N + (j)) * (v)->G + (i)]
#define A2(arr, j, i) (arr)[(size_t)(j) * G + (i)]

/* src/odevec.F90:127-134 */
static void set_coeff(ovec_t* v) {
  static const double c[6][7] = {
      {1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      {2.0 / 3.0, 1.0, 1.0 / 3.0, 0.0, 0.0, 0.0, 0.0},
      {6.0 / 11.0, 1.0, 6.0 / 11.0, 1.0 / 11.0, 0.0, 0.0, 0.0},
      {12.0 / 25.0, 1.0, 7.0 / 10.0, 1.0 / 5.0, 1.0 / 50.0, 0.0, 0.0},
      {60.0 / 137.0, 1.0, 225.0 / 274.0, 85.0 / 274.0, 15.0 / 274.0, 1.0 / 274.0, 0.0},
      {20.0 / 49.0, 1.0, 58.0 / 63.0, 5.0 / 12.0, 25.0 / 252.0, 1.0 / 84.0, 1.0 / 1764.0}};
  memset(v->coeff, 0, sizeof v->coeff);
  for (int q = 1; q <= 6; ++q)
    for (int kk = 0; kk < 7; ++kk) v->coeff[q][kk] = c[q - 1][kk];
}

void oracle_coeff_table(double* out /* [6*7] */) {
  ovec_t v;
  set_coeff(&v);
  for (int q = 1; q <= 6; ++q)
    for (int kk = 0; kk < 7; ++kk) out[(q - 1) * 7 + kk] = v.coeff[q][kk];
}

/* src/odevec.F90:1034-1054 */
static double tau(int order, int order2, const ovec_t* v) {
  double fac = 1.0;
  for (int i = 1; i <= order; ++i) fac = fac * i;
  if (order - 1 == order2) return (order2 + 1.0) / fac;
  return (order2 + 1.0) / (fac * v->coeff[order][order]);
}

static int active(const ovec_t* v, int i) { return !v->mask || v->mask[i]; }

/* WeightedNorm1 (src/odevec.F90:669-684): max over (masked) cells */
static double wnorm(const ovec_t* v, const double* x) {
  const int G = v->G, N = v->N;
  double m = -INFINITY;
  int nonnan = 0, nan = 0;
  for (int i = 0; i < G; ++i) {
    if (!active(v, i)) continue;
    double s = (A2(x, 0, i) * A2(x, 0, i)) * A2(v->w, 0, i);
    for (int j = 1; j < N; ++j) s = s + (A2(x, j, i) * A2(x, j, i)) * A2(v->w, j, i);
    const double r = sqrt(s / (double)N);
    if (r != r) { nan = 1; continue; }
    nonnan = 1;
    if (r > m) m = r;
  }
  return (nan && !nonnan) ? NAN : m;
}

static double wnorm_col(const ovec_t* v, int col) {
  return wnorm(v, v->yNS + (size_t)col * v->N * v->G);
}

/* CalcErrorWeightInvSquared (src/odevec.F90:706-716) */
static void calc_weights(ovec_t* v, const octl_t* c, const double* y) {
  const size_t M = (size_t)v->N * v->G;
  for (size_t e = 0; e < M; ++e) {
    const double r = 1.0 / (c->rtol * fabs(y[e]) + c->atol);
    v->w[e] = r * r;
  }
}

static void rhs_all(ovec_t* v, const double* y, double* f) {
  const int G = v->G, N = v->N;
  double *yc = v->tmp, *fc = v->tmp + N;
  for (int i = 0; i < G; ++i) {
    for (int j = 0; j < N; ++j) yc[j] = A2(y, j, i);
    oracle_rhs(v->n, yc, v->k + (size_t)i * (v->n->nrate > 0 ? v->n->nrate : 1), fc);
    for (int j = 0; j < N; ++j) A2(f, j, i) = fc[j];
  }
}

/* PredictSolution (src/odevec.F90:769-793) */
static void predict(ovec_t* v) {
  const size_t M = (size_t)v->N * v->G;
  for (int l = 0; l <= v->order - 1; ++l)
    for (int kk = v->order; kk >= l + 1; --kk) {
      double* a = v->yNS + (size_t)(kk - 1) * M;
      const double* b = v->yNS + (size_t)kk * M;
      for (size_t e = 0; e < M; ++e) a[e] = b[e] + a[e];
    }
  memset(v->en, 0, M * sizeof(double));
}

/* ResetSystem (src/odevec.F90:526-544) */
static void reset_system(ovec_t* v) {
  const size_t M = (size_t)v->N * v->G;
  for (int kk = 0; kk <= v->order - 1; ++kk)
    for (int j = v->order; j >= kk + 1; --j) {
      double* a = v->yNS + (size_t)(j - 1) * M;
      const double* b = v->yNS + (size_t)j * M;
      for (size_t e = 0; e < M; ++e) a[e] = a[e] - b[e];
    }
}

/* SetStepSize (src/odevec.F90:720-743) */
static void set_step(ovec_t* v, double scale, double* dt) {
  const size_t M = (size_t)v->N * v->G;
  double d = 1.0;
  for (int kk = 1; kk <= v->order; ++kk) {
    d = scale * d;
    double* a = v->yNS + (size_t)kk * M;
    for (size_t e = 0; e < M; ++e) a[e] = a[e] * d;
  }
  *dt = *dt * scale;
  v->jchange = v->jchange * scale;
}

/* CalcStepSizeOrder (src/odevec.F90:547-618); returns 0 or -1 on order error */
static int step_order(ovec_t* v, double* scale) {
  const int q = v->order;
  const size_t M = (size_t)v->N * v->G;
  double down, same, up;
  if (q == 1) down = 0.0;
  else down = 1.0 / (F32(1.3) * o_pow(wnorm_col(v, q) / tau(q, q - 1, v), 1.0 / q) + 1e-6);
  same = 1.0 / (F32(1.2) * o_pow(wnorm(v, v->en) / tau(q, q, v), 1.0 / (q + 1.0)) + 1e-6);
  if (q == MAXORD || v->neglect_up) up = 0.0;
  else {
    for (size_t e = 0; e < M; ++e) v->den[e] = v->en[e] - v->enold[e];
    up = 1.0 / (F32(1.4) * o_pow(wnorm(v, v->den) / tau(q, q + 1, v), 1.0 / (q + 2.0)) + 1e-6);
  }
  int loc = 1;
  double best = down;
  if (same > best) { loc = 2; best = same; }
  if (up > best) { loc = 3; best = up; }
  double s = best;
  if (s >= F32(1.1)) {
    if (loc == 1) {
      if (q == 1) return -1;
      v->order = q - 1;
    } else if (loc == 2) {
      s = same;
    } else {
      if (q == MAXORD) return -1;
      double* a = v->yNS + (size_t)(q + 1) * M;
      for (size_t e = 0; e < M; ++e) a[e] = (v->coeff[q][q] * v->en[e]) / (q + 1.0);
      v->order = q + 1;
    }
  } else {
    s = same;
  }
  const double cap = v->first_step ? 1e4 : 10.0;
  v->first_step = 0;
  s = fmin(cap, s);
  v->jchange = v->jchange * s;
  v->swu = 0;
  v->neglect_up = 0;
  *scale = s;
  return 0;
}

/* CalcJac_dense (src/odevec.F90:469-506), LUmethod 3: P = I - dt*beta*J with
 * column k of J by a forward difference of the RHS in species k, step
 * r = max(srur*|y_k|, r0/sqrt(w_k)), r0 = 1e3*|dt|*uround*neq*||f|| (the
 * weighted norm over the vector, masked cells excluded). Writes P of every
 * cell, row-major, to Pall[i*N*N ...]. The literals uround = 2^-52 and srur =
 * 2^-26 are exact in the reference's default real. */
static void fd_p_all(ovec_t* v, const double* y, double dt, double beta, double* Pall) {
  const int G = v->G, N = v->N;
  const double uround = 2.2204460492503131e-16, srur = 1.4901161193847656e-08;
  double* yy = (double*)malloc(sizeof(double) * (size_t)N * G);
  double* f0 = (double*)malloc(sizeof(double) * (size_t)N * G);
  double* D = (double*)malloc(sizeof(double) * (size_t)N * G);
  double* r = (double*)malloc(sizeof(double) * (size_t)G);
  double* fac = (double*)malloc(sizeof(double) * (size_t)G);
  memcpy(yy, y, sizeof(double) * (size_t)N * G);
  rhs_all(v, yy, f0);
  const double nrm = wnorm(v, f0);
  double r0 = 1e3 * fabs(dt) * uround * (double)N * nrm;
  if (r0 == 0.0) r0 = 1.0;
  for (int k = 0; k < N; ++k) {
    for (int i = 0; i < G; ++i) {
      const double ytmp = A2(yy, k, i);
      r[i] = fmax(srur * fabs(ytmp), r0 / sqrt(A2(v->w, k, i)));
      A2(yy, k, i) = ytmp + r[i];
      fac[i] = -(dt * beta / r[i]);
    }
    rhs_all(v, yy, D);
    for (int i = 0; i < G; ++i) {
      for (int j = 0; j < N; ++j) Pall[(size_t)i * N * N + (size_t)j * N + k] = (A2(D, j, i) - A2(f0, j, i)) * fac[i];
      A2(yy, k, i) = A2(y, k, i);
    }
  }
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < N; ++j) Pall[(size_t)i * N * N + (size_t)j * N + j] = 1.0 + Pall[(size_t)i * N * N + (size_t)j * N + j];
  free(yy); free(f0); free(D); free(r); free(fac);
}

/* GetLU (src/odevec_commons.F90:55-72, entries written by pre_odevec.py:132-139),
 * LUmethod 2: the dense factors of one cell by the closed-form program, each
 * operation one correctly rounded IEEE operation in program order; entries
 * outside the outputs are zero. A(m,j) at A[m*N+j]. */
static void closed_lu(const onet_t* n, const double* y, const double* k, double hb, double* A) {
  const int N = n->neq;
  double* v = (double*)malloc(sizeof(double) * (size_t)(n->lc_nops > 0 ? n->lc_nops : 1));
  for (int i = 0; i < n->lc_nops; ++i) {
    const int a = n->lc_a[i], b = n->lc_b[i];
    switch (n->lc_code[i]) {
      case 0: v[i] = y[a]; break;
      case 1: v[i] = k[a]; break;
      case 2: v[i] = hb; break;
      case 3: v[i] = n->lc_c[i]; break;
      case 4: v[i] = v[a] + v[b]; break;
      case 5: v[i] = v[a] - v[b]; break;
      case 6: v[i] = v[a] * v[b]; break;
      case 7: v[i] = v[a] / v[b]; break;
      default: v[i] = -v[a]; break;
    }
  }
  memset(A, 0, sizeof(double) * (size_t)N * N);
  for (int o = 0; o < n->lc_nout; ++o) {
    const int* e = n->lc_out + 3 * o;
    A[e[0] * N + e[1]] = v[e[2]];
  }
  free(v);
}

/* GetJac (or CalcJac, LUmethod 3) + LUDecompose for all cells; GetLU for
 * LUmethod 2 */
static void factor_all(ovec_t* v, const double* y, double dt) {
  const onet_t* n = v->n;
  const int G = v->G, N = v->N;
  const double hb = v->coeff[v->order][0] * dt;
  double* yc = v->tmp;
  double* P = v->tmp + 2 * N;
  double* Pfd = NULL;
  if (n->lu_method == 3) {
    Pfd = (double*)malloc(sizeof(double) * (size_t)N * N * G);
    fd_p_all(v, y, dt, v->coeff[v->order][0], Pfd);
  }
  for (int i = 0; i < G; ++i) {
    for (int j = 0; j < N; ++j) yc[j] = A2(y, j, i);
    if (n->lu_method == 2 && !n->csc) {
      closed_lu(n, yc, v->k + (size_t)i * (n->nrate > 0 ? n->nrate : 1), hb, v->LU + (size_t)i * N * N);
      continue;
    }
    if (Pfd) memcpy(P, Pfd + (size_t)i * N * N, sizeof(double) * (size_t)N * N);
    else oracle_p_dense(n, yc, v->k + (size_t)i * (n->nrate > 0 ? n->nrate : 1), hb, P);
    double* A = v->LU + (size_t)i * (n->csc ? n->nnz : N * N);
    if (n->csc) {
      /* jac%sdata in CSC order (fill entries are structural zeros of P) */
      for (int col = 0; col < N; ++col)
        for (int kk = n->u_col_start[col]; kk < n->u_col_start[col + 1]; ++kk)
          A[kk] = P[n->row_index[kk] * N + col];
      /* LUDecompose_sparse (src/odevec.F90:929-971) with a dense work column */
      double* w = v->tmp + 2 * N + N * N;
      for (int col = 0; col < N; ++col) {
        for (int kk = n->u_col_start[col]; kk < n->u_col_start[col + 1]; ++kk) w[n->row_index[kk]] = A[kk];
        for (int kk = n->u_col_start[col]; kk < n->l_col_start[col]; ++kk) {
          const int j = n->row_index[kk];
          const double alpha = w[j] / A[n->l_col_start[j]];
          w[j] = alpha;
          for (int jj = n->l_col_start[j] + 1; jj < n->u_col_start[j + 1]; ++jj)
            w[n->row_index[jj]] = w[n->row_index[jj]] - alpha * A[jj];
        }
        for (int kk = n->u_col_start[col]; kk < n->u_col_start[col + 1]; ++kk) A[kk] = w[n->row_index[kk]];
      }
    } else {
      /* LUDecompose_dense (src/odevec.F90:898-922), A(m,j) at A[m*N+j] */
      memcpy(A, P, (size_t)N * N * sizeof(double));
      for (int kk = 0; kk < N - 1; ++kk) {
        for (int j = kk + 1; j < N; ++j) A[j * N + kk] = A[j * N + kk] / A[kk * N + kk];
        for (int j = kk + 1; j < N; ++j)
          for (int m = kk + 1; m < N; ++m) A[m * N + j] = A[m * N + j] - A[m * N + kk] * A[kk * N + j];
      }
    }
  }
  free(Pfd);
}

/* permute residual, SolveLU, unpermute (src/odevec.F90:328-342, 797-891) */
static void solve_all(ovec_t* v) {
  const onet_t* n = v->n;
  const int G = v->G, N = v->N;
  double* x = v->tmp;
  double* t = v->tmp + N;
  for (int i = 0; i < G; ++i) {
    for (int j = 0; j < N; ++j) x[j] = A2(v->res, n->perm[j], i);
    const double* A = v->LU + (size_t)i * (n->csc ? n->nnz : N * N);
    if (n->csc) {
      for (int kk = 0; kk < N; ++kk) {
        const double mult = x[kk] / A[n->l_col_start[kk]];
        x[kk] = mult;
        for (int q = n->l_col_start[kk] + 1; q < n->u_col_start[kk + 1]; ++q)
          x[n->row_index[q]] = x[n->row_index[q]] - mult * A[q];
      }
      for (int kk = N - 1; kk >= 0; --kk) {
        const double mult = x[kk];
        for (int q = n->u_col_start[kk]; q < n->l_col_start[kk]; ++q)
          x[n->row_index[q]] = x[n->row_index[q]] - mult * A[q];
      }
    } else {
      for (int kk = 0; kk < N; ++kk) {
        t[kk] = 0.0;
        for (int j = 0; j < kk; ++j) t[kk] = t[kk] + A[kk * N + j] * t[j];
        t[kk] = x[kk] - t[kk];
      }
      for (int kk = N - 1; kk >= 0; --kk) {
        x[kk] = 0.0;
        for (int j = kk + 1; j < N; ++j) x[kk] = x[kk] + A[kk * N + j] * x[j];
        x[kk] = t[kk] - x[kk];
        x[kk] = x[kk] / A[kk * N + kk];
      }
    }
    for (int j = 0; j < N; ++j) A2(v->den, n->perm[j], i) = x[j];
  }
}

static int any_neg(const ovec_t* v, const double* a) {
  const int G = v->G;
  for (int i = 0; i < G; ++i) {
    if (!active(v, i)) continue;
    for (int j = 0; j < v->N; ++j)
      if (A2(a, j, i) < 0.0) return 1;
  }
  return 0;
}

static void copy_masked(const ovec_t* v, double* dst, const double* src) {
  const int G = v->G;
  for (int i = 0; i < G; ++i) {
    if (!active(v, i)) continue;
    for (int j = 0; j < v->N; ++j) A2(dst, j, i) = A2(src, j, i);
  }
}

enum { OST_OK = 0, OST_DT_MIN = 1, OST_CONV_FAIL = 2, OST_ORDER = 3, OST_MAX_STEPS = 4, OST_SKIPPED = 5 };

/* SolveLinearSystem (src/odevec.F90:242-465); returns a status or -1 to continue */
static int solve_linear_system(ovec_t* v, const octl_t* c, double* t, double* dt, double* y, ocnt_t* cnt) {
  const int G = v->G, N = v->N;
  const size_t M = (size_t)N * G;
  const double told = *t;
  for (;;) { /* predictor loop */
    cnt->attempts++;
    cnt->qsum += v->order;
    if (cnt->attempts > c->max_attempts) return OST_MAX_STEPS;
    *t = *t + *dt;
    predict(v);
    if (c->check_negatives && any_neg(v, v->yNS)) {
      *t = told;
      reset_system(v);
      set_step(v, 0.25, dt);
      if (*dt <= c->dt_min) return OST_DT_MIN;
      continue;
    }
    copy_masked(v, y, v->yNS);
    const double beta = v->coeff[v->order][0];
    if (v->upd_jac) {
      factor_all(v, y, *dt);
      v->upd_jac = 0;
      v->lju = 0;
      v->rate = 0.7;
      v->jchange = 1.0;
      cnt->jac++;
    }
    int failed = 0, converged = 0, it = 0;
    const double tq = tau(v->order, v->order, v);
    while (!converged) {
      rhs_all(v, y, v->rhs);
      /* CalcResiduum (src/odevec.F90:975-990) */
      for (size_t e = 0; e < M; ++e) v->res[e] = *dt * v->rhs[e] - (v->yNS[M + e] + v->en[e]);
      solve_all(v);
      for (size_t e = 0; e < M; ++e) v->en[e] = v->en[e] + v->den[e];
      for (int i = 0; i < G; ++i) {
        if (!active(v, i)) continue;
        for (int j = 0; j < N; ++j) A2(y, j, i) = YNS(v, 0, j, i) + beta * A2(v->en, j, i);
      }
      cnt->newton++;
      /* CheckConvergence (src/odevec.F90:622-665) */
      const double e = wnorm(v, v->den);
      double rtmp = 0.0;
      if (it != 0) {
        rtmp = e / v->conv_err;
        v->rate = fmax(0.2 * v->rate, rtmp);
      }
      const double crit = e * fmin(1.0, 1.5 * v->rate) / (tq / (2.0 * (v->order + 2.0)));
      int nan = 0;
      for (size_t q = 0; q < M; ++q) nan |= (v->den[q] != v->den[q]) || (v->w[q] != v->w[q]);
      if (nan) failed = 1;
      else if (crit < 1.0) converged = 1;
      else if (it >= 2 && v->rate > 2.0 * rtmp) failed = 1;
      else if (it == 3) failed = 1;
      else { failed = 0; v->conv_err = e; it++; }
      if (c->check_negatives && any_neg(v, y)) { failed = 1; converged = 0; }
      if (failed && !converged) break;
    }
    if (failed && !converged) {
      *t = told;
      v->fc++;
      cnt->convfail++;
      if (v->fc > 10) return OST_CONV_FAIL;
      reset_system(v);
      set_step(v, 0.25, dt);
      if (*dt < c->dt_min) return OST_DT_MIN;
      v->upd_jac = 1;
      continue;
    }
    v->fc = 0;
    const double err = wnorm(v, v->en) / tq;
    calc_weights(v, c, y);
    if (err > 1.0) {
      cnt->errfail++;
      if (v->fet >= 3) {
        *t = told;
        v->order = 1;
        copy_masked(v, y, v->yNS);
        rhs_all(v, y, v->rhs);
        for (size_t q = 0; q < M; ++q) v->yNS[M + q] = *dt * v->rhs[q];
        v->upd_jac = 1;
      } else {
        *t = told;
        reset_system(v);
        v->upd_jac = 1;
        v->swu = 0;
        v->neglect_up = 1;
      }
      double s;
      cnt->stepsel++;
      if (step_order(v, &s)) return OST_ORDER;
      set_step(v, s, dt);
      v->fet++;
      continue;
    }
    v->swu++;
    v->lju++;
    /* UpdateNordsieck (src/odevec.F90:747-765) */
    for (int kk = 0; kk <= v->order; ++kk) {
      const double lk = v->coeff[v->order][kk];
      double* a = v->yNS + (size_t)kk * M;
      for (size_t q = 0; q < M; ++q) a[q] = a[q] + v->en[q] * lk;
    }
    if (v->swu == v->order + 1) {
      double s;
      cnt->stepsel++;
      if (step_order(v, &s)) return OST_ORDER;
      set_step(v, s, dt);
    }
    v->fet = 0;
    cnt->steps++;
    break;
  }
  if (v->lju >= 20) v->upd_jac = 1;
  else if (fabs(v->jchange - 1.0) > F32(0.3)) v->upd_jac = 1;
  memcpy(v->enold, v->en, M * sizeof(double));
  return -1;
}

/*
 * SolveODE (src/odevec.F90:158-214) for one vector of G cells.
 * y [N][G] in/out, prm [nparam][G] (may be NULL for nparam 0), mask [G] or NULL.
 * *time, *dt in/out like the reference. Returns a status.
 * Deviation: t_stop <= time leaves y unchanged (the reference divides by a
 * zero step size when t_stop == time and extrapolates backwards otherwise).
 */
int oracle_solve_vector(const onet_t* n, int G, double* time, double* dt, double t_stop, double* y,
                        const double* prm, const char* mask, const octl_t* c, ocnt_t* cnt) {
  const int N = n->neq;
  const size_t M = (size_t)N * G;
  const int nr = n->nrate > 0 ? n->nrate : 1;
  const int lusz = n->csc ? n->nnz : N * N;
  ovec_t v;
  memset(&v, 0, sizeof v);
  memset(cnt, 0, sizeof *cnt);
  v.n = n;
  v.G = G;
  v.N = N;
  v.mask = mask;
  set_coeff(&v);
  double* mem = (double*)calloc((size_t)(MAXORD + 2) * M + 8 * M + (size_t)lusz * G + (size_t)nr * G +
                                    3 * (size_t)N + (size_t)N * N + 16,
                                sizeof(double));
  if (!mem) return -100;
  double* p = mem;
  v.yNS = p; p += (MAXORD + 2) * M;
  v.en = p; p += M;
  v.enold = p; p += M;
  v.den = p; p += M;
  v.w = p; p += M;
  v.rhs = p; p += M;
  v.res = p; p += M;
  v.LU = p; p += (size_t)lusz * G;
  v.k = p; p += (size_t)nr * G;
  v.tmp = p;
  int any = 0;
  for (int i = 0; i < G; ++i) any |= active(&v, i);
  if (!any || !(*time < t_stop)) {
    *time = t_stop;
    free(mem);
    return OST_SKIPPED;
  }
  /* rates once per cell from its initial state */
  {
    double* yc = v.tmp;
    double pc[n->nparam > 0 ? n->nparam : 1];
    for (int i = 0; i < G; ++i) {
      for (int j = 0; j < N; ++j) yc[j] = A2(y, j, i);
      for (int q = 0; q < n->nparam; ++q) pc[q] = prm[(size_t)q * G + i];
      cell_rates(n, yc, pc, v.k + (size_t)i * nr);
    }
  }
  v.first_step = 1;
  v.order = 1;
  v.swu = 0;
  v.rate = 0.7;
  v.jchange = 0.0;
  v.neglect_up = 0;
  v.fet = 0;
  v.fc = 0;
  v.upd_jac = 1;
  calc_weights(&v, c, y);
  memcpy(v.yNS, y, M * sizeof(double));
  rhs_all(&v, y, v.rhs);
  for (size_t e = 0; e < M; ++e) v.yNS[M + e] = *dt * v.rhs[e];
  /* CalcInitialStepSize (src/odevec.F90:216-238) */
  {
    const double tol = c->rtol;
    /* maxval over the cells: NaN sums are skipped unless all are NaN (as in wnorm) */
    double mx = -INFINITY;
    int nonnan = 0, nan = 0;
    for (int i = 0; i < G; ++i) {
      if (!active(&v, i)) continue;
      double s = (A2(v.rhs, 0, i) * A2(v.rhs, 0, i)) * ((A2(v.w, 0, i) * tol) * tol);
      for (int j = 1; j < N; ++j) s = s + (A2(v.rhs, j, i) * A2(v.rhs, j, i)) * ((A2(v.w, j, i) * tol) * tol);
      if (s != s) { nan = 1; continue; }
      nonnan = 1;
      if (s > mx) mx = s;
    }
    if (nan && !nonnan) mx = NAN;
    const double W0 = fmax(fabs(*time), fabs(t_stop));
    double d = sqrt(tol / (o_pow(W0, -2.0) + mx / (double)N));
    *dt = fmin(d, fabs(t_stop - *time));
  }
  int st = OST_OK;
  while (*time < t_stop) {
    const int r = solve_linear_system(&v, c, time, dt, y, cnt);
    if (r >= 0) { st = r; break; }
  }
  if (st == OST_OK) {
    /* InterpolateSolution (src/odevec.F90:998-1030) */
    const int q = v.order;
    const double rel = (t_stop - *time) / *dt;
    for (int i = 0; i < G; ++i) {
      if (!active(&v, i)) continue;
      for (int j = 0; j < N; ++j) {
        double val = YNS(&v, q, j, i);
        for (int kk = 1; kk <= q; ++kk) val = YNS(&v, q - kk, j, i) + rel * val;
        A2(y, j, i) = val;
      }
    }
    *time = t_stop;
  } else {
    /* the reference STOPs here; the solution at the failure is y_NS(:,:,0) */
    copy_masked(&v, y, v.yNS);
  }
  free(mem);
  return st;
}

/*
 * Per-cell batch: cell i integrates alone (G = 1) from t0[i] to tstop[i].
 * y [N][ncells], prm [nparam][ncells]; status/dt_out/cnt_out optional.
 * Threads: nthreads > 0 uses OpenMP when built with it.
 */
void oracle_solve_cells(const onet_t* n, long ncells, const double* t0, const double* tstop,
                        const double* dt0, double* y, const double* prm, const octl_t* c, int* status,
                        double* dt_out, int* cnt_out, int nthreads) {
  const int N = n->neq;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
  for (long i = 0; i < ncells; ++i) {
    double yc[N], pc[n->nparam > 0 ? n->nparam : 1];
    for (int j = 0; j < N; ++j) yc[j] = y[(size_t)j * ncells + i];
    for (int q = 0; q < n->nparam; ++q) pc[q] = prm[(size_t)q * ncells + i];
    double t = t0[i], dt = dt0[i];
    ocnt_t cnt;
    int st = oracle_solve_vector(n, 1, &t, &dt, tstop[i], yc, pc, NULL, c, &cnt);
    for (int j = 0; j < N; ++j) y[(size_t)j * ncells + i] = yc[j];
    if (status) status[i] = st;
    if (dt_out) dt_out[i] = dt;
    if (cnt_out) {
      const int* cc = &cnt.attempts;
      for (int m = 0; m < 8; ++m) cnt_out[(size_t)m * ncells + i] = cc[m];
    }
  }
}

/* LU factor of one dense P (network pattern) for tests: returns the factors
 * in the network's storage (CSC order or dense row-major). */
void oracle_factor_one(const onet_t* n, const double* y, const double* k, double hb, double* out) {
  ovec_t v;
  memset(&v, 0, sizeof v);
  const int N = n->neq;
  v.n = n;
  v.G = 1;
  v.N = N;
  set_coeff(&v);
  v.order = 1;
  double* tmp = (double*)calloc((size_t)4 * N + (size_t)N * N + 16, sizeof(double));
  v.tmp = tmp;
  v.k = (double*)k;
  v.LU = out;
  /* factor_all uses coeff[order][0]*dt as hb: order 1 has beta = 1 */
  factor_all(&v, y, hb);
  free(tmp);
}

/* CalcJac_dense's P for one cell with weights w, step dt and beta, for tests */
void oracle_fd_p_one(const onet_t* n, const double* y, const double* k, const double* w, double dt, double beta,
                     double* P) {
  ovec_t v;
  memset(&v, 0, sizeof v);
  const int N = n->neq;
  v.n = n;
  v.G = 1;
  v.N = N;
  double* tmp = (double*)calloc((size_t)4 * N + 16, sizeof(double));
  v.tmp = tmp;
  v.k = (double*)k;
  v.w = (double*)w;
  fd_p_all(&v, y, dt, beta, P);
  free(tmp);
}

/* one permuted solve with factors `lu` (network storage) for tests */
void oracle_solve_one(const onet_t* n, const double* lu, const double* res, double* den) {
  ovec_t v;
  memset(&v, 0, sizeof v);
  const int N = n->neq;
  v.n = n;
  v.G = 1;
  v.N = N;
  double* tmp = (double*)calloc((size_t)4 * N + 16, sizeof(double));
  v.tmp = tmp;
  v.LU = (double*)lu;
  v.res = (double*)res;
  v.den = den;
  solve_all(&v);
  free(tmp);
}

void oracle_rates_one(const onet_t* n, const double* y0, const double* p, double* k) {
  cell_rates(n, y0, p, k);
}

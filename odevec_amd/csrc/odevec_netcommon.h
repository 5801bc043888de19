// Shared helpers for the generated network headers: host/device qualifiers,
// integer powers and the KROME primordial rate coefficients.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "odevec_math.h"

#define ODV_HD __host__ __device__ __attribute__((always_inline))

// Kernels and integrator templates live in a variant namespace so that the
// default build (FMA contraction on) and the reproducible build (compiled
// with -ffp-contract=off -DODV_REPRO, bit-identical to the oracle) can be
// linked into one library.
#ifdef ODV_REPRO
#define ODV_VARIANT_NS repro
#define ODV_REPRO_FLAG 1
#else
#define ODV_VARIANT_NS fast
#define ODV_REPRO_FLAG 0
#endif

// x**n for a constant integer n >= 1 by right-to-left binary exponentiation
// (x^2 = x*x, x^3 = x*(x*x), x^4 = (x*x)*(x*x), ...). Every evaluator of the
// rate formulas in this project uses this same product order.
ODV_HD inline double odv_ipow(double x, int n) {
#pragma clang fp contract(off)
  double r = 1.0;
  bool have = false;
  double b = x;
  while (n) {
    if (n & 1) {
      r = have ? r * b : b;
      have = true;
    }
    n >>= 1;
    if (n) b = b * b;
  }
  return r;
}

// c[0] + c[1]*x + c[2]*x^2 + ... summed left to right, powers via odv_ipow.
template <int N>
ODV_HD inline double odv_lnpoly(const double (&c)[N], double x) {
#pragma clang fp contract(off)
  double s = c[0];
#pragma unroll
  for (int i = 1; i < N; ++i) s = s + c[i] * odv_ipow(x, i);
  return s;
}

// Fortran default-real (single precision) literals promoted to double.
#define ODV_F32(v) ((double)(v##f))

// Rate coefficients of the 16-species primordial network, k[0..37]
// (src/odevec_commons.F90:101-384). Each rate is small + fit(T); a rate that
// is piecewise in T is just small outside its range (the reference initialises
// every rate to small and fills each range with a masked assignment,
// src/odevec_commons.F90:139-141). Tgas is read from the state slot `tgas`
// (the network reserves that slot, pre_odevec/example_systems.py:61); the
// reference fixes Tgas = 1e3 K for every cell (src/odevec_commons.F90:120-124).
// small = 1e-40 / max(max_j y_j, 1)^3 (src/odevec_commons.F90:127-130) is taken
// from the state passed in. Elementary functions are the deterministic ones
// of odevec_math.h and nothing is contracted, so every build of this function
// returns the same bits.
ODV_HD inline void odv_primordial_rates(const double* y, int neq, int tgas, double* k) {
#pragma clang fp contract(off)
  double nmax = y[0];
  for (int j = 1; j < neq; ++j) nmax = fmax(nmax, y[j]);
  nmax = fmax(nmax, 1.0);
  const double small = 1e-40 / (nmax * nmax * nmax);
  const double T = y[tgas];
  const double Te = T * 8.617343e-5;
  const double lnTe = odv::m_log(Te);
  const double T32 = T * ODV_F32(0.0033333333333333335);
  const double invT = 1.0 / T;
  const double invTe = 1.0 / Te;
  const double sqrTgas = sqrt(T);
  for (int i = 0; i < 38; ++i) k[i] = small;

  {  // H + E -> H+ + E + E
    const double c[9] = {-32.71396786, 13.5365560, -5.73932875, 1.56315498, -0.28770560,
                         3.48255977e-2, -2.63197617e-3, 1.11954395e-4, -2.03914985e-6};
    k[0] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  if (T <= 5.5e3) k[1] = small + 3.92e-13 * odv::m_pow(invTe, 0.6353);  // H+ + E -> H
  if (T > 5.5e3) {                                               // H+ + E -> H
    const double c[10] = {-28.61303380689232, -0.7241125657826851, -0.02026044731984691,
                          -0.002380861877349834, -0.0003212605213188796, -0.00001421502914054107,
                          4.989108920299513e-6, 5.755614137575758e-7, -1.856767039775261e-8,
                          -3.071135243196595e-9};
    k[2] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  {  // HE + E -> HE+ + E + E
    const double c[9] = {-44.09864886, 23.91596563, -10.7532302, 3.05803875, -0.56851189,
                         6.79539123e-2, -5.00905610e-3, 2.06723616e-4, -3.64916141e-6};
    k[3] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  if (T <= 9.28e3) k[4] = small + 3.92e-13 * odv::m_pow(invTe, 0.6353);  // HE+ + E -> HE
  if (T > 9.28e3)                                                 // HE+ + E -> HE
    k[5] = small + (1.54e-9 * (1.0 + 0.3 / odv::m_exp(8.099328789667 * invTe)) /
                        (odv::m_exp(40.49664394833662 * invTe) * odv::m_pow(Te, 1.5)) +
                    3.92e-13 / odv::m_pow(Te, 0.6353));
  {  // HE+ + E -> HE++ + E + E
    const double c[9] = {-68.71040990212001, 43.93347632635, -18.48066993568, 4.701626486759002,
                         -0.7692466334492, 0.08113042097303, -0.005324020628287001,
                         0.0001975705312221, -3.165581065665e-6};
    k[6] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  // HE++ + E -> HE+
  k[7] = small + (3.36e-10 / sqrTgas / odv::m_pow(T / 1.e3, 0.2) / (1.0 + odv::m_pow(T / 1.e6, 0.7)));
  // H + E -> H-
  k[8] = small + 6.77e-15 * odv::m_pow(Te, 0.8779);
  if (T < 1160.0) k[9] = small + 1.43e-9;  // H- + H -> H2 + E
  if (T > 1160.0) {                         // H- + H -> H2 + E
    const double c[8] = {-20.06913897587003, 0.2289800603272916, 0.03599837721023835,
                         -0.004555120027032095, -0.0003105115447124016, 0.0001073294010367247,
                         -8.36671960467864e-6, 2.238306228891639e-7};
    k[10] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  if (T <= 6.7e3) k[11] = small + 1.85e-23 * odv::m_pow(T, 1.8);  // H + H+ -> H2+
  if (T > 6.7e3)                                            // H + H+ -> H2+
    k[12] = small + 5.81e-16 * odv::m_pow(T / 5.62e4, -0.6657 * odv::m_log10(T / 5.62e4));
  k[13] = small + 6.0e-10;  // H2+ + H -> H2 + H+
  if (T > 3.48e3) {         // H2 + H+ -> H2+ + H
    const double c[9] = {-24.24914687731536, 3.400824447095291, -3.898003964650152,
                         2.045587822403071, -0.5416182856220388, 0.0841077503763412,
                         -0.007879026154483455, 0.0004138398421504563, -9.36345888928611e-6};
    k[14] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  // H2 + E -> H + H + E
  k[15] = small + 5.6e-11 * odv::m_exp(-102124.0 * invT) * odv::m_pow(T, 0.5);
  // H2 + H -> H + H + H
  k[16] = small + 1.0670825e-10 * odv::m_pow(Te, 2.012) * odv::m_exp(-4.463 * invTe) / odv::m_pow(1.0 + 0.2472 * Te, 3.512);
  {  // H- + E -> H + E + E
    const double c[9] = {-18.01849334273, 2.360852208681, -0.2827443061704, 0.01623316639567,
                         -0.03365012031362999, 0.01178329782711, -0.001656194699504,
                         0.0001068275202678, -2.631285809207e-6};
    k[17] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  if (T <= 1.16e3) k[18] = small + 2.56e-9 * odv::m_pow(Te, 1.78186);  // H- + H -> H + H + E
  if (T > 1.16e3) {                                              // H- + H -> H + H + E
    const double c[10] = {-20.37260896533324, 1.139449335841631, -0.1421013521554148,
                          0.00846445538663, -0.0014327641212992, 0.0002012250284791,
                          0.0000866396324309, -0.00002585009680264, 2.4555011970392e-6,
                          -8.06838246118e-8};
    k[19] = small + odv::m_exp(odv_lnpoly(c, lnTe));
  }
  k[20] = small + 6.5e-9 / sqrt(Te);            // H- + H+ -> H + H
  k[21] = small + 1.e-8 * odv::m_pow(T, -0.4);         // H- + H+ -> H2+ + E
  if (T <= 6.17e2) k[22] = small + 1.e-8;       // H2+ + E -> H + H
  if (T > 6.17e2) k[23] = small + 1.32e-6 * odv::m_pow(T, -0.76);
  k[24] = small + 5.e-7 * sqrt(1.e2 * invT);    // H2+ + H- -> H + H2
  if (T <= 3e2) k[25] = small + 1.3e-32 * odv::m_pow(T32, -0.38);        // H + H + H -> H2 + H
  if (T > 3e2) k[26] = small + 1.3e-32 * odv::m_pow(T32, -1.00);
  if (T <= 3e2) k[27] = small + 1.3e-32 * odv::m_pow(T32, -0.38) / 8.0;  // H2 + H + H -> H2 + H2
  if (T > 3e2) k[28] = small + 1.3e-32 * odv::m_pow(T32, -1.00) / 8.0;
  if (T >= 5e1)  // H+ + D -> H + D+
    k[29] = small + (2.00e-10 * odv::m_pow(T, 0.402) * odv::m_exp(-37.1 * invT) - 3.31e-17 * odv::m_pow(T, 1.48));
  if (T >= 5e1)  // H + D+ -> H+ + D
    k[30] = small + (2.06e-10 * odv::m_pow(T, ODV_F32(0.396)) * odv::m_exp(-33.0 * invT) +
                     2.03e-9 * odv::m_pow(T, -ODV_F32(0.332)));
  {  // H2 + D+ -> HD + H+
    const double l = odv::m_log10(T);
    k[31] = small + 1.e-9 * (ODV_F32(0.417) + ODV_F32(0.846) * l - ODV_F32(0.137) * odv_ipow(l, 2));
  }
  k[32] = small + 1.0e-9 * odv::m_exp(-4.57e2 * invT);  // HD + H+ -> H2 + D+
  if (T <= 2.e3) {                               // H2 + D -> HD + H
    const double l = odv::m_log10(T);
    k[33] = small + odv::m_pow(10.0, -ODV_F32(56.4737) + ODV_F32(5.88886) * l + ODV_F32(7.19692) * odv_ipow(l, 2) +
                                  ODV_F32(2.25069) * odv_ipow(l, 3) - ODV_F32(2.16903) * odv_ipow(l, 4) +
                                  ODV_F32(0.317887) * odv_ipow(l, 5));
  }
  if (T > 2e3) k[34] = small + 3.17e-10 * odv::m_exp(-5207.0 * invT);  // H2 + D -> HD + H
  if (T > 2e2)                                                    // HD + H -> H2 + D
    k[35] = small + 5.25e-11 * odv::m_exp(-4430.0 * invT + 1.739e5 * odv_ipow(invT, 2));
  k[36] = small + 1.5e-9 * odv::m_pow(T32, -0.1);             // D + H- -> HD + E
  k[37] = small + 3.6e-12 * odv::m_pow(T / 300.0, -0.75);     // D+ + E -> D
}

// Device Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt for gfx950.
//
// Two evaluations:
//   * exact (production): F_{M}(T) from a Taylor expansion about the nearest
//     grid point of a table (spacing 1/16, 8 terms, |error| < 3e-17 relative),
//     then downward recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1), which is
//     stable.  For T >= 36 the asymptote F_0 = sqrt(pi/T)/2 (error < e^{-T})
//     and the upward recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T) are used.
//     The table is filled on the host in long double from the all-positive
//     series and lives in device memory (111 KB FP64, read through L1/L2),
//     followed by exp(-k/16) for the same grid.
//   * compat (parity with the reference): the Numerical-Recipes incomplete
//     gamma of libpyquante2 (cints.c:302-373) evaluated per m with its own
//     constants (EPS = 3e-7, ITMAX = 100, x clamped to >= 1e-8), so integrals
//     reproduce the reference's Boys truncation error exactly.  Used only
//     where bit-level agreement with the reference's integrals is requested.
#pragma once
#include <hip/hip_runtime.h>

namespace rchem {

constexpr int kBoysCols = 24;          // m = 0 .. 23 stored per grid point
constexpr int kBoysTaylor = 8;         // Taylor terms
constexpr int kBoysMaxM = kBoysCols - kBoysTaylor;  // highest m served: 16
constexpr double kBoysTMax = 36.0;
constexpr double kBoysInvH = 16.0;     // grid spacing 1/16
constexpr int kBoysRows = 36 * 16 + 2; // T in [0, 36] rounded to the grid, plus margin

enum BoysMode { kBoysExact = 0, kBoysCompat = 1 };

// 1/sqrt(x) and 1/x from the hardware estimates refined by Newton steps
// (two in FP64, one in FP32): ~1 ulp, a fraction of the cost of the IEEE
// division and square root sequences.
__device__ __forceinline__ double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * __builtin_fma(-h * y, y, 1.5);
  y = y * __builtin_fma(-h * y, y, 1.5);
  return y;
}
__device__ __forceinline__ float fast_rsqrt(float x) {
  float y = __builtin_amdgcn_rsqf(x);
  y = y * __builtin_fmaf(-0.5f * x * y, y, 1.5f);
  return y;
}

// Exact Boys F_0..F_M into F[0..M].  `table` holds the Boys grid followed by
// exp(-k/16) for the same k; e^{-T} = e^{-T_k} e^{T_k - T} with an 8-term
// Taylor series of the second factor (|T_k - T| <= 1/32).
template <int M, typename R>
__device__ __forceinline__ void boys_exact(R T, const R* __restrict__ table, R* __restrict__ F) {
  if (T < R(kBoysTMax)) {
    const int k = static_cast<int>(T * R(kBoysInvH) + R(0.5));
    const R d = R(k) * (R(1) / R(kBoysInvH)) - T;  // T0 - T
    const R* __restrict__ row = table + k * kBoysCols + M;
    // F_M(T) = sum_j F_{M+j}(T0) (T0 - T)^j / j!
    R f = row[7];
    f = row[6] + d * (R(1.0 / 7.0) * f);
    f = row[5] + d * (R(1.0 / 6.0) * f);
    f = row[4] + d * (R(1.0 / 5.0) * f);
    f = row[3] + d * (R(1.0 / 4.0) * f);
    f = row[2] + d * (R(1.0 / 3.0) * f);
    f = row[1] + d * (R(1.0 / 2.0) * f);
    f = row[0] + d * f;
    F[M] = f;
    if constexpr (M > 0) {
      R p = R(1.0 / 5040.0);
      p = R(1.0 / 720.0) + d * p;
      p = R(1.0 / 120.0) + d * p;
      p = R(1.0 / 24.0) + d * p;
      p = R(1.0 / 6.0) + d * p;
      p = R(0.5) + d * p;
      p = R(1) + d * p;
      p = R(1) + d * p;
      const R e = table[kBoysRows * kBoysCols + k] * p;
      const R t2 = T + T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * R(1.0 / (2 * m + 1));
    }
  } else {
    const R rt = fast_rsqrt(T);
    F[0] = R(0.88622692545275801365) * rt;  // sqrt(pi)/2 / sqrt(T); erf(sqrt T) = 1 - O(e^-T)
    if constexpr (M > 0) {
      // e^{-T} < 2.4e-16 here: it only corrects the upward recursion, so a
      // single-precision exponential is accurate enough
      const R e = static_cast<R>(__expf(-static_cast<float>(T)));
      const R h = R(0.5) * rt * rt;  // 1/(2T)
#pragma unroll
      for (int m = 0; m < M; ++m) F[m + 1] = (R(2 * m + 1) * F[m] - e) * h;
    }
  }
}

// Compat: the reference's Fgamma(m, x) (cints.c:302-373), NR gser/gcf.
__device__ __forceinline__ double fgamma_compat(double m, double x) {
  const double EPS = 3.0e-7, FPMIN = 1.0e-30;
  if (fabs(x) < 1.0e-8) x = 1.0e-8;
  const double a = m + 0.5;
  const double gln = lgamma(a);
  double val;
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    val = __builtin_nan("");
    for (int n = 1; n <= 100; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * EPS) {
        const double gamser = sum * exp(-x + a * log(x) - gln);
        val = exp(gln) * gamser;
        break;
      }
    }
  } else {
    double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
    for (int i = 1; i <= 100; ++i) {
      const double an = -i * (i - a);
      b += 2.0;
      d = an * d + b;
      if (fabs(d) < FPMIN) d = FPMIN;
      c = b + an / c;
      if (fabs(c) < FPMIN) c = FPMIN;
      d = 1.0 / d;
      const double del = d * c;
      h *= del;
      if (fabs(del - 1.0) < EPS) break;
    }
    const double gammcf = exp(-x + a * log(x) - gln) * h;
    val = exp(gln) * (1.0 - gammcf);
  }
  return 0.5 * pow(x, -m - 0.5) * val;
}

template <int M>
__device__ __forceinline__ void boys_compat(double T, double* __restrict__ F) {
#pragma unroll
  for (int m = 0; m <= M; ++m) F[m] = fgamma_compat(double(m), T);
}

template <int M, int MODE, typename R>
__device__ __forceinline__ void boys(R T, const R* __restrict__ table, R* __restrict__ F) {
  if constexpr (MODE == kBoysCompat) {
    boys_compat<M>(T, F);
  } else {
    boys_exact<M, R>(T, table, F);
  }
}

}  // namespace rchem

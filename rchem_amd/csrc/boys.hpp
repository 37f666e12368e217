// Device Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt for gfx950.
//
// Two evaluations:
//   * exact (production): F_{M}(T) from a Taylor expansion about the nearest
//     grid point T0 of a table (spacing 1/8, |T - T0| <= 1/16; 9 terms in FP64,
//     truncation (1/16)^9/9! < 5e-17 relative, 6 in FP32, < 3e-10), then downward recursion
//     F_m = (2T F_{m+1} + e^{-T}) / (2m+1), which is stable, with
//     e^{-T} = e^{-T0} e^{T0 - T} from the same row and a Taylor series.  For
//     T >= 36 the asymptote F_0 = sqrt(pi/T)/2 (error < e^{-T}) and the upward
//     recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T) are used.  The Taylor
//     coefficients F_{M+j}(T0)/j! are tabulated per order M (so the Horner
//     steps are bare FMAs), filled on the host in long double from the
//     all-positive series; a kernel stages its order's block in LDS, so every
//     lookup is an LDS read, not a cache miss.
//   * compat (parity with the reference): the Numerical-Recipes incomplete
//     gamma of libpyquante2 (cints.c:302-373) evaluated per m with its own
//     constants (EPS = 3e-7, ITMAX = 100, x clamped to >= 1e-8), so integrals
//     reproduce the reference's Boys truncation error exactly.  Used only
//     where bit-level agreement with the reference's integrals is requested.
#pragma once
#include <hip/hip_runtime.h>

namespace rchem {

constexpr int kBoysMaxM = 9;                          // highest m served
// The finer the grid, the fewer Taylor terms a lookup reads from LDS: at
// spacing 1/8 a lookup reads 9 doubles (72 B) where spacing 1/2 needed 13
// (104 B); the Horner chains are shorter by the same count.
constexpr int kBoysTerms = 9;                         // Taylor terms, FP64
constexpr int kBoysTermsF = 6;                        // Taylor terms, FP32
constexpr double kBoysTMax = 36.0;
constexpr double kBoysInvH = 8.0;                     // grid spacing 1/8
constexpr int kBoysRows = 289;                        // T0 = 0, 1/8, ..., 36
// Device layout: for M = 0..kBoysMaxM a block of kBoysRows rows holding
// F_{M+j}(T0) / j!, j = 0..kBoysTerms-1 (the Taylor coefficients of F_M about
// T0 in powers of T0 - T), then one row of exp(-T0).  A kernel of order M
// stages its own block (20.8 KB in FP64) and, for M >= 2 (downward
// recursion), the exp row (2.3 KB).
constexpr int kBoysBlock = kBoysRows * kBoysTerms;
constexpr int kBoysExpOffset = (kBoysMaxM + 1) * kBoysBlock;
constexpr int kBoysTableSize = kBoysExpOffset + kBoysRows;
constexpr int kBoysStageSize = kBoysBlock + kBoysRows;
// Staged entries a kernel of highest order M needs (no exp row for M <= 1).
__host__ __device__ constexpr int boys_stage_size(int M) {
  return M <= 1 ? kBoysBlock : kBoysStageSize;
}

enum BoysMode { kBoysExact = 0, kBoysCompat = 1 };

// 1/sqrt(x) from the hardware estimate.  FP64: the estimate (measured on
// gfx950: 5.2e-8 relative over [0.05, 2e4]) refined by one third-order
// (Halley) step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, whose error is
// O(e^3) -- five operations in a chain of five, where two Newton steps take
// six in a chain of seven; ~1 ulp, a fraction of the cost of the IEEE
// division and square root sequences.  FP32: the estimate as it is, within
// 0.86 ulp (measured on gfx950 over [1e-3, 1e5] against FP64); one Newton
// step in FP32 (four operations) measured 1.53 ulp, its own rounding
// outweighing what it corrects.
__device__ __forceinline__ double fast_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
}
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// Nearest grid point of T (0 <= T < kBoysTMax): its index k and d = T0 - T,
// |d| <= 1/16.  Rounding by adding the power of two whose ulp is the grid
// spacing (2^52 h in FP64, 2^23 h in FP32): the sum r = kMagic + T0 has k in
// its low mantissa bits (read off the register, no conversion), r - kMagic is
// T0 exactly, and T0 - T is exact (a multiple of T's ulp below 1/16).  Three
// adds with the constant in a scalar register, where a float-to-int and an
// int-to-float conversion (or an FMA against a constant that must be rebuilt
// in vector registers, gfx9 VOP3 taking no 64-bit literal) cost more.  A tie
// rounds to even, which keeps |d| <= 1/16.
static_assert(kBoysInvH == 8.0, "boys_grid's constants assume spacing 1/8");
template <typename R>
__device__ __forceinline__ void boys_grid(R T, int* k, R* d) {
  if constexpr (sizeof(R) == 8) {
    constexpr double kMagic = 562949953421312.0;  // 2^49: ulp 1/8 on [2^49, 2^50)
    const double r = T + kMagic;
    *k = static_cast<int>(static_cast<unsigned>(__double_as_longlong(r)));
    *d = (r - kMagic) - T;
  } else {
    constexpr float kMagic = 1048576.0f;  // 2^20: ulp 1/8 on [2^20, 2^21)
    const float r = T + kMagic;
    *k = static_cast<int>(__float_as_uint(r) & 0x7fffffu);
    *d = (r - kMagic) - T;
  }
}

// Order of the table block a kernel of highest order M stages: for M = 1 the
// order-0 block, whose Taylor polynomial and its derivative give F_0 and F_1
// in one pass (dF_0/dT = -F_1), with no exponential.
constexpr int boys_block_order(int M) { return M == 1 ? 0 : M; }

// Exact Boys F_0..F_M into F[0..M]; `table` is the staged block of order
// boys_block_order(M) (kBoysBlock coefficients, then the exp row), normally in LDS.
template <int M, typename R>
__device__ __forceinline__ void boys_exact(R T, const R* __restrict__ table, R* __restrict__ F) {
  static_assert(M <= kBoysMaxM, "Boys order beyond the table");
  constexpr int K = sizeof(R) == 8 ? kBoysTerms : kBoysTermsF;
  if (T < R(kBoysTMax)) {
    int k;
    R d;  // T0 - T, |d| <= 1/16
    boys_grid<R>(T, &k, &d);
    // 0 <= k < kBoysRows: a 24-bit multiply (full rate; a 32-bit one is
    // quarter rate)
    const R* __restrict__ c = table + __umul24(static_cast<unsigned>(k), kBoysTerms);
    if constexpr (M == 1) {
      // F_0(T0 - d) = sum_j c_j d^j (c_j = F_j(T0) / j!); F_1 = d/dd of it.
      // The derivative's chain starts at c_{K-1} (its first step, 0 d + p,
      // is not left to the compiler, which may not drop a product with 0)
      R q = c[K - 1];
      R p = q * d + c[K - 2];
#pragma unroll
      for (int j = K - 3; j >= 0; --j) {
        q = q * d + p;
        p = p * d + c[j];
      }
      F[0] = p;
      F[1] = q;
      return;
    }
    R f = c[K - 1];
#pragma unroll
    for (int j = K - 2; j >= 0; --j) f = c[j] + d * f;
    F[M] = f;
    if constexpr (M > 0 && sizeof(R) == 4) {
      const R e = __expf(-T);  // hardware exp2: ample for the FP32 path
      const R t2 = T + T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * R(1.0 / (2 * m + 1));
    } else if constexpr (M > 0) {
      // e^{d} by its Taylor series with constant coefficients 1/j!
      static_assert(K <= 14, "inv_fact holds 14 terms");
      constexpr double inv_fact[14] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                       1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800,
                                       1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0};
      R p = R(inv_fact[K - 1]);
#pragma unroll
      for (int j = K - 2; j >= 0; --j) p = R(inv_fact[j]) + d * p;
      const R e = table[kBoysBlock + k] * p;  // e^{-T0} e^{d}
      const R t2 = T + T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * R(1.0 / (2 * m + 1));
    }
  } else {
    const R rt = fast_rsqrt(T);
    F[0] = R(0.88622692545275801365) * rt;  // sqrt(pi)/2 / sqrt(T); erf(sqrt T) = 1 - O(e^-T)
    if constexpr (M == 1) {
      // F_1 = (F_0 - e^{-T}) / (2T): e^{-T} / F_0 < 2e-15 for T >= 36, so the
      // exponential is dropped (the branch costs a few operations per lane)
      F[1] = F[0] * (R(0.5) * rt * rt);
    } else if constexpr (M > 1) {
      // e^{-T} < 2.4e-16 here: it only corrects the upward recursion, so a
      // single-precision exponential is accurate enough
      const R e = static_cast<R>(__expf(-static_cast<float>(T)));
      const R h = R(0.5) * rt * rt;  // 1/(2T)
#pragma unroll
      for (int m = 0; m < M; ++m) F[m + 1] = (R(2 * m + 1) * F[m] - e) * h;
    }
  }
}

// Orders 0 and 1 without a branch: the table form (at the clamped grid point
// for T beyond the table) and the asymptotic form are both evaluated and the
// one that applies is selected, so the caller's loop body stays one basic
// block.  Higher orders take boys_exact.
template <int M, typename R>
__device__ __forceinline__ void boys_exact_select(R T, const R* __restrict__ table,
                                                  R* __restrict__ F) {
  if constexpr (M > 1) {
    boys_exact<M, R>(T, table, F);
  } else {
    constexpr int K = sizeof(R) == 8 ? kBoysTerms : kBoysTermsF;
    const bool near = T < R(kBoysTMax);
    const R Tn = near ? T : R(0);
    int k;
    R d;
    boys_grid<R>(Tn, &k, &d);
    const R* __restrict__ c = table + __umul24(static_cast<unsigned>(k), kBoysTerms);
    const R rt = fast_rsqrt(near ? R(1) : T);
    const R f0 = R(0.88622692545275801365) * rt;
    if constexpr (M == 1) {
      R q = c[K - 1];
      R p = q * d + c[K - 2];
#pragma unroll
      for (int j = K - 3; j >= 0; --j) {
        q = q * d + p;
        p = p * d + c[j];
      }
      F[0] = near ? p : f0;
      F[1] = near ? q : f0 * (R(0.5) * rt * rt);
    } else {
      R p = c[K - 1];
#pragma unroll
      for (int j = K - 2; j >= 0; --j) p = p * d + c[j];
      F[0] = near ? p : f0;
    }
  }
}

// Runtime-order version over the full device table (one-electron integrals).
__device__ __forceinline__ void boys_exact_rt(int M, double T, const double* __restrict__ table,
                                              double* F) {
  if (T < kBoysTMax) {
    const int k = static_cast<int>(T * kBoysInvH + 0.5);
    const double d = k * (1.0 / kBoysInvH) - T;
    const double* c = table + M * kBoysBlock + k * kBoysTerms;
    double f = c[kBoysTerms - 1];
    for (int j = kBoysTerms - 2; j >= 0; --j) f = c[j] + d * f;
    F[M] = f;
    double p = 1.0;
    for (int j = kBoysTerms - 1; j >= 1; --j) p = 1.0 + d * (p / j);
    const double e = table[kBoysExpOffset + k] * p;
    for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
  } else {
    F[0] = 0.88622692545275801365 / sqrt(T);
    const double e = exp(-T);
    for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

// Copies the block of order boys_block_order(M) and, where the recursion needs
// it (M >= 2), the exp row into LDS: boys_stage_size(M) entries (every thread of
// the workgroup calls it; a barrier follows).
template <int M, typename R>
__device__ __forceinline__ void stage_boys(R* __restrict__ dst, const R* __restrict__ src) {
  const int m = boys_block_order(M <= kBoysMaxM ? M : kBoysMaxM);
  for (int k = threadIdx.x; k < kBoysBlock; k += blockDim.x) dst[k] = src[m * kBoysBlock + k];
  if constexpr (boys_stage_size(M) > kBoysBlock)
    for (int k = threadIdx.x; k < kBoysRows; k += blockDim.x)
      dst[kBoysBlock + k] = src[kBoysExpOffset + k];
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

template <int M, int MODE, typename R, bool SELECT = false>
__device__ __forceinline__ void boys(R T, const R* __restrict__ table, R* __restrict__ F) {
  if constexpr (MODE == kBoysCompat) {
    boys_compat<M>(T, F);
  } else if constexpr (SELECT) {
    boys_exact_select<M, R>(T, table, F);
  } else {
    boys_exact<M, R>(T, table, F);
  }
}

}  // namespace rchem

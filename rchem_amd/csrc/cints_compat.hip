// Primitive one-electron integrals and the Boys function behind the cints.h
// compatibility exports (rchem_amd/csrc/cints_compat.cpp), evaluated on the
// GPU, one thread per call:
//   overlap  (pi/g)^(3/2) exp(-a1 a2 |AB|^2 / g) prod_x s_x with the 1D overlap
//            s = sum_i f_2i(l1, l2, PA, PB) (2i-1)!! / (2g)^i (binomial
//            expansion of the product of the two polynomial factors);
//   kinetic  from overlaps with the second function's powers shifted by +-2;
//   nuclear  McMurchie--Davidson: -(2 pi / g) exp(...) sum E^x_t E^y_u E^z_v
//            R_tuv(g, P - C), with the reference's incomplete-gamma Boys function;
//   Fgamma   that Boys function itself (cints.c:302-373 semantics).
// The reference functions these serve are cints.c:152-243 and cints.c:302.
#include <hip/hip_runtime.h>

#include "boys.hpp"

namespace rchem {

struct CompatRec {
  int kind;  // 0 overlap, 1 kinetic, 2 nuclear attraction (unit charge, no norms), 3 Fgamma
  int l1[3], l2[3];
  double a1, a2;
  double A[3], B[3], C[3];
  double m, x;  // Fgamma arguments
  double out;
};

namespace {

constexpr int kCompatL = 8;  // highest power per axis of the overlap / kinetic functions
constexpr int kNucL = 8;     // highest total angular momentum l1 + l2 of nuclear attraction

__device__ double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

__device__ double dfact_odd(int n) {  // n!! for odd n >= -1
  double r = 1.0;
  for (int i = n; i > 1; i -= 2) r *= i;
  return r;
}

// coefficient of x^s in (x + pa)^ia (x + pb)^ib
__device__ double binomial_coef(int s, int ia, int ib, double pa, double pb) {
  double sum = 0.0;
  for (int t = 0; t <= s; ++t) {
    const int i = s - t, j = t;  // powers of x taken from each factor
    if (i > ia || j > ib) continue;
    sum += binom(ia, i) * binom(ib, j) * pow(pa, ia - i) * pow(pb, ib - j);
  }
  return sum;
}

__device__ double overlap_1d(int l1, int l2, double pa, double pb, double g) {
  double sum = 0.0;
  for (int i = 0; 2 * i <= l1 + l2; ++i)
    sum += binomial_coef(2 * i, l1, l2, pa, pb) * dfact_odd(2 * i - 1) / pow(2.0 * g, i);
  return sum;
}

__device__ double overlap_prim(double a1, const int* l1, const double* A, double a2, const int* l2,
                               const double* B) {
  const double g = a1 + a2;
  double ab2 = 0.0, s = 1.0;
  for (int x = 0; x < 3; ++x) {
    ab2 += (A[x] - B[x]) * (A[x] - B[x]);
    const double p = (a1 * A[x] + a2 * B[x]) / g;
    s *= overlap_1d(l1[x], l2[x], p - A[x], p - B[x], g);
  }
  return pow(3.14159265358979323846 / g, 1.5) * exp(-a1 * a2 * ab2 / g) * s;
}

__device__ double kinetic_prim(double a1, const int* l1, const double* A, double a2, const int* l2,
                               const double* B) {
  const int L2 = l2[0] + l2[1] + l2[2];
  double t = a2 * (2 * L2 + 3) * overlap_prim(a1, l1, A, a2, l2, B);
  for (int x = 0; x < 3; ++x) {
    int p[3] = {l2[0], l2[1], l2[2]};
    p[x] += 2;
    t -= 2.0 * a2 * a2 * overlap_prim(a1, l1, A, a2, p, B);
    if (l2[x] >= 2) {
      int m[3] = {l2[0], l2[1], l2[2]};
      m[x] -= 2;
      t -= 0.5 * l2[x] * (l2[x] - 1) * overlap_prim(a1, l1, A, a2, m, B);
    }
  }
  return t;
}

// Hermite expansion E_t of x_A^i x_B^j for one axis (t = 0 .. i + j)
__device__ void hermite_1d(int i, int j, double pa, double pb, double oo2p, double* E) {
  double cur[kNucL + 1], nxt[kNucL + 1];
  for (int t = 0; t <= i + j; ++t) cur[t] = 0.0;
  cur[0] = 1.0;
  int deg = 0;
  for (int k = 0; k < i + j; ++k) {
    const double shift = k < i ? pa : pb;
    for (int t = 0; t <= deg + 1; ++t) {
      double v = shift * (t <= deg ? cur[t] : 0.0);
      if (t > 0) v += oo2p * cur[t - 1];
      if (t + 1 <= deg) v += (t + 1) * cur[t + 1];
      nxt[t] = v;
    }
    ++deg;
    for (int t = 0; t <= deg; ++t) cur[t] = nxt[t];
  }
  for (int t = 0; t <= i + j; ++t) E[t] = cur[t];
}

__device__ double nuclear_prim(double a1, const int* l1, const double* A, double a2, const int* l2,
                               const double* B, const double* C) {
  const double g = a1 + a2, oog = 1.0 / g;
  double ab2 = 0.0, P[3], PC[3];
  for (int x = 0; x < 3; ++x) {
    ab2 += (A[x] - B[x]) * (A[x] - B[x]);
    P[x] = (a1 * A[x] + a2 * B[x]) * oog;
    PC[x] = P[x] - C[x];
  }
  double Ex[kNucL + 1], Ey[kNucL + 1], Ez[kNucL + 1];
  hermite_1d(l1[0], l2[0], P[0] - A[0], P[0] - B[0], 0.5 * oog, Ex);
  hermite_1d(l1[1], l2[1], P[1] - A[1], P[1] - B[1], 0.5 * oog, Ey);
  hermite_1d(l1[2], l2[2], P[2] - A[2], P[2] - B[2], 0.5 * oog, Ez);
  const int tx = l1[0] + l2[0], ty = l1[1] + l2[1], tz = l1[2] + l2[2];
  const int L = tx + ty + tz;
  const double T = g * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  // R_{n,tuv}: level n from level n + 1, kept as two flat levels
  constexpr int D = kNucL + 1;
  double Rn[D * D * D], Rc[D * D * D];
  double m2p = 1.0;
  for (int k = 0; k < L; ++k) m2p *= -2.0 * g;
  Rn[0] = m2p * fgamma_compat(static_cast<double>(L), T);
  for (int lev = L - 1; lev >= 0; --lev) {
    m2p /= -2.0 * g;
    const int top = L - lev;
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          double r;
          if (t + u + v == 0) {
            r = m2p * fgamma_compat(static_cast<double>(lev), T);
          } else if (t > 0) {
            r = PC[0] * Rn[((t - 1) * D + u) * D + v];
            if (t > 1) r += (t - 1) * Rn[((t - 2) * D + u) * D + v];
          } else if (u > 0) {
            r = PC[1] * Rn[(t * D + u - 1) * D + v];
            if (u > 1) r += (u - 1) * Rn[(t * D + u - 2) * D + v];
          } else {
            r = PC[2] * Rn[(t * D + u) * D + v - 1];
            if (v > 1) r += (v - 1) * Rn[(t * D + u) * D + v - 2];
          }
          Rc[(t * D + u) * D + v] = r;
        }
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) Rn[(t * D + u) * D + v] = Rc[(t * D + u) * D + v];
  }
  double sum = 0.0;
  for (int t = 0; t <= tx; ++t)
    for (int u = 0; u <= ty; ++u)
      for (int v = 0; v <= tz; ++v) sum += Ex[t] * Ey[u] * Ez[v] * Rn[(t * D + u) * D + v];
  return -2.0 * 3.14159265358979323846 * oog * exp(-a1 * a2 * ab2 * oog) * sum;
}

__global__ void compat_kernel(CompatRec* recs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CompatRec& r = recs[i];
  double v = 0.0;
  switch (r.kind) {
    case 0: v = overlap_prim(r.a1, r.l1, r.A, r.a2, r.l2, r.B); break;
    case 1: v = kinetic_prim(r.a1, r.l1, r.A, r.a2, r.l2, r.B); break;
    case 2: v = nuclear_prim(r.a1, r.l1, r.A, r.a2, r.l2, r.B, r.C); break;
    default: v = fgamma_compat(r.m, r.x); break;
  }
  r.out = v;
}

}  // namespace

// Largest power per axis the kernel serves (per function; kinetic adds 2).
int compat_max_power() { return kCompatL - 2; }
// Largest total angular momentum (both functions) of nuclear attraction.
int compat_max_nuclear_l() { return kNucL; }

hipError_t compat_launch(CompatRec* recs, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(compat_kernel, dim3((n + 63) / 64), dim3(64), 0, s, recs, n);
  return hipGetLastError();
}

}  // namespace rchem

// One-electron integrals S, T, V over contracted Cartesian Gaussians (gfx950).
//
// One thread per shell pair (a >= b) writes the full component blocks of S, T
// and V and their transposes.  McMurchie--Davidson: per axis the Hermite
// expansion coefficients E_t^{ij}, overlap S_x(i,j) = E_0^{ij} sqrt(pi/p),
// kinetic energy from overlaps with b shifted by +-2 (the identity rchem's
// C library uses, cints.c:152-177), nuclear attraction
//   <a|-1/r_C|b> = -(2 pi / p) sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P - C)
// with the exact Boys function (rchem's V goes through boys::micb25,
// os86.rs:770).  Normalisation and coefficient conventions are rchem's
// (basis.rs:140-149, 201): the exponent-dependent norm is folded into the
// contraction coefficients on the host, the per-component factor is applied
// here.  Matrix assembly mirrors S/T/V in basis.rs:214-338.
#include <hip/hip_runtime.h>

#include "boys.hpp"

namespace rchem {

struct ShellDev {
  double xyz[3];
  int l, nprim, p0, f0;  // angular momentum, primitives, first primitive, first function
};

struct OneElArgs {
  const ShellDev* shells;
  const double* exps;
  const double* ccoef;  // c_i * radial norm
  const int2* pairs;    // (a, b), a >= b
  int npairs;
  int nbf;
  int natom;
  const double* atom_xyz;
  const double* atom_z;
  const double* boys_table;
  double* S;
  double* T;
  double* V;
};

constexpr int kLmax1 = 2;              // highest l served
constexpr int kE = kLmax1 + 3;         // E index range for j up to lb + 2
constexpr int kT = 2 * kLmax1 + 3;     // Hermite t range

__device__ __forceinline__ int ncart1(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian component k of shell l in rchem's order (shell.rs:1-12).
__device__ __forceinline__ void cart_comp(int l, int k, int* c) {
  int idx = 0;
  for (int a = 1; a <= l + 1; ++a)
    for (int b = 1; b <= a; ++b) {
      if (idx == k) {
        c[0] = l + 1 - a;
        c[1] = a - b;
        c[2] = b - 1;
        return;
      }
      ++idx;
    }
}

__device__ __forceinline__ double comp_factor(const int* c) {
  // 1/sqrt((2l-1)!!(2m-1)!!(2n-1)!!), values for l <= 2 per axis: 1, 1, 1/sqrt(3)
  double f = 1.0;
  for (int k = 0; k < 3; ++k)
    if (c[k] == 2) f *= 0.57735026918962576451;
  return f;
}

// E_t^{ij} for i <= imax, j <= jmax, one axis, without the exp(-mu X^2) factor.
__device__ __forceinline__ void hermite_e(int imax, int jmax, double pa, double pb, double oo2p,
                                          double E[kLmax1 + 1][kE][kT]) {
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      for (int t = 0; t < kT; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
  for (int i = 0; i <= imax; ++i) {
    if (i > 0) {
      for (int t = 0; t <= i; ++t) {
        double v = pa * E[i - 1][0][t];
        if (t > 0) v += oo2p * E[i - 1][0][t - 1];
        if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
        E[i][0][t] = v;
      }
    }
    for (int j = 1; j <= jmax; ++j)
      for (int t = 0; t <= i + j; ++t) {
        double v = pb * E[i][j - 1][t];
        if (t > 0) v += oo2p * E[i][j - 1][t - 1];
        if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        E[i][j][t] = v;
      }
  }
}

__global__ __launch_bounds__(128) void one_electron_kernel(const OneElArgs a) {
  const int pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= a.npairs) return;
  const int2 pr = a.pairs[pid];
  const ShellDev sa = a.shells[pr.x], sb = a.shells[pr.y];
  const int la = sa.l, lb = sb.l;
  const int na = ncart1(la), nb = ncart1(lb);
  constexpr int NMAX = 6;
  double s[NMAX][NMAX], t[NMAX][NMAX], v[NMAX][NMAX];
  for (int i = 0; i < NMAX; ++i)
    for (int j = 0; j < NMAX; ++j) s[i][j] = t[i][j] = v[i][j] = 0.0;
  int ca[NMAX][3], cb[NMAX][3];
  for (int k = 0; k < na; ++k) cart_comp(la, k, ca[k]);
  for (int k = 0; k < nb; ++k) cart_comp(lb, k, cb[k]);
  const double AB2 = (sa.xyz[0] - sb.xyz[0]) * (sa.xyz[0] - sb.xyz[0]) +
                     (sa.xyz[1] - sb.xyz[1]) * (sa.xyz[1] - sb.xyz[1]) +
                     (sa.xyz[2] - sb.xyz[2]) * (sa.xyz[2] - sb.xyz[2]);

  double E[3][kLmax1 + 1][kE][kT];
  for (int ip = 0; ip < sa.nprim; ++ip)
    for (int jp = 0; jp < sb.nprim; ++jp) {
      const double al = a.exps[sa.p0 + ip], be = a.exps[sb.p0 + jp];
      const double p = al + be, oop = 1.0 / p;
      const double cc = a.ccoef[sa.p0 + ip] * a.ccoef[sb.p0 + jp];
      const double kab = exp(-al * be * oop * AB2);
      double P[3];
      for (int x = 0; x < 3; ++x) {
        P[x] = (al * sa.xyz[x] + be * sb.xyz[x]) * oop;
        hermite_e(la, lb + 2, P[x] - sa.xyz[x], P[x] - sb.xyz[x], 0.5 * oop, E[x]);
      }
      const double sq = sqrt(3.14159265358979323846 * oop);
      const double pre_s = cc * kab * sq * sq * sq;
      // overlap and kinetic
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
          const int* A = ca[i];
          const int* B = cb[j];
          double ov[3], ovp2[3], ovm2[3];
          for (int x = 0; x < 3; ++x) {
            ov[x] = E[x][A[x]][B[x]][0];
            ovp2[x] = E[x][A[x]][B[x] + 2][0];
            ovm2[x] = B[x] >= 2 ? E[x][A[x]][B[x] - 2][0] : 0.0;
          }
          const double s0 = ov[0] * ov[1] * ov[2];
          s[i][j] += pre_s * s0;
          const int lbt = B[0] + B[1] + B[2];
          double t1 = ovp2[0] * ov[1] * ov[2] + ov[0] * ovp2[1] * ov[2] + ov[0] * ov[1] * ovp2[2];
          double t2 = B[0] * (B[0] - 1) * ovm2[0] * ov[1] * ov[2] +
                      B[1] * (B[1] - 1) * ov[0] * ovm2[1] * ov[2] +
                      B[2] * (B[2] - 1) * ov[0] * ov[1] * ovm2[2];
          t[i][j] += pre_s * (be * (2 * lbt + 3) * s0 - 2.0 * be * be * t1 - 0.5 * t2);
        }
      // nuclear attraction: R_tuv via the Hermite recursion, one centre at a time
      const int L = la + lb;
      for (int c = 0; c < a.natom; ++c) {
        const double PC[3] = {P[0] - a.atom_xyz[3 * c], P[1] - a.atom_xyz[3 * c + 1],
                              P[2] - a.atom_xyz[3 * c + 2]};
        const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double F[2 * kLmax1 + 1];
        boys_exact_rt(L, T, a.boys_table, F);
        // R[n][t][u][v], t+u+v <= L - n
        double R[2 * kLmax1 + 1][2 * kLmax1 + 1][2 * kLmax1 + 1][2 * kLmax1 + 1];
        double m2p = 1.0;
        for (int n = 0; n <= L; ++n) {
          R[n][0][0][0] = m2p * F[n];
          m2p *= -2.0 * p;
        }
        for (int n = L - 1; n >= 0; --n) {
          const int top = L - n;
          for (int tt = 0; tt <= top; ++tt)
            for (int uu = 0; uu <= top - tt; ++uu)
              for (int vv = 0; vv <= top - tt - uu; ++vv) {
                if (tt + uu + vv == 0) continue;
                double r;
                if (tt > 0) {
                  r = PC[0] * R[n + 1][tt - 1][uu][vv];
                  if (tt > 1) r += (tt - 1) * R[n + 1][tt - 2][uu][vv];
                } else if (uu > 0) {
                  r = PC[1] * R[n + 1][tt][uu - 1][vv];
                  if (uu > 1) r += (uu - 1) * R[n + 1][tt][uu - 2][vv];
                } else {
                  r = PC[2] * R[n + 1][tt][uu][vv - 1];
                  if (vv > 1) r += (vv - 1) * R[n + 1][tt][uu][vv - 2];
                }
                R[n][tt][uu][vv] = r;
              }
        }
        const double pre_v = -a.atom_z[c] * cc * kab * 2.0 * 3.14159265358979323846 * oop;
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            const int* A = ca[i];
            const int* B = cb[j];
            double sum = 0.0;
            for (int tt = 0; tt <= A[0] + B[0]; ++tt)
              for (int uu = 0; uu <= A[1] + B[1]; ++uu)
                for (int vv = 0; vv <= A[2] + B[2]; ++vv)
                  sum += E[0][A[0]][B[0]][tt] * E[1][A[1]][B[1]][uu] * E[2][A[2]][B[2]][vv] *
                         R[0][tt][uu][vv];
            v[i][j] += pre_v * sum;
          }
      }
    }
  const size_t n = static_cast<size_t>(a.nbf);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      const double f = comp_factor(ca[i]) * comp_factor(cb[j]);
      const size_t mu = sa.f0 + i, nu = sb.f0 + j;
      if (a.S) { a.S[mu * n + nu] = f * s[i][j]; a.S[nu * n + mu] = f * s[i][j]; }
      if (a.T) { a.T[mu * n + nu] = f * t[i][j]; a.T[nu * n + mu] = f * t[i][j]; }
      if (a.V) { a.V[mu * n + nu] = f * v[i][j]; a.V[nu * n + mu] = f * v[i][j]; }
    }
}

hipError_t one_electron_launch(const OneElArgs& a, hipStream_t s) {
  if (a.npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(one_electron_kernel, dim3((a.npairs + 127) / 128), dim3(128), 0, s, a);
  return hipGetLastError();
}

}  // namespace rchem

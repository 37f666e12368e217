// One-electron integrals S, T, V over contracted Cartesian Gaussians (gfx950).
//
// One thread per shell pair (a >= b), one kernel per (la, lb) class, writes the
// full component blocks of S, T and V and their transposes.  McMurchie--Davidson: per axis the Hermite
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
  const int* shell_mol;   // several molecules: the molecule of each shell, else null
  const int2* mol_atoms;  // nuclei [begin, end) of each molecule
  int cls_begin[10];      // pairs sorted by class 3 la + lb: class c is [cls_begin[c], cls_begin[c + 1])
};

constexpr int kLmax1 = 2;              // highest l served
constexpr int kE = kLmax1 + 3;         // E index range for j up to lb + 2

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

// Hermite coefficients E_t^{ij}, i <= IMAX, j <= JMAX, one axis, with
// compile-time bounds so that every index is static (registers, no scratch).
template <int IMAX, int JMAX>
__device__ __forceinline__ void hermite_e_t(double pa, double pb, double oo2p,
                                            double (&E)[IMAX + 1][JMAX + 1][IMAX + JMAX + 1]) {
#pragma unroll
  for (int i = 0; i <= IMAX; ++i)
#pragma unroll
    for (int j = 0; j <= JMAX; ++j)
#pragma unroll
      for (int t = 0; t <= IMAX + JMAX; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
#pragma unroll
  for (int i = 0; i <= IMAX; ++i) {
    if (i > 0) {
#pragma unroll
      for (int t = 0; t <= i; ++t) {
        double v = pa * E[i - 1][0][t];
        if (t > 0) v += oo2p * E[i - 1][0][t - 1];
        if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
        E[i][0][t] = v;
      }
    }
#pragma unroll
    for (int j = 1; j <= JMAX; ++j)
#pragma unroll
      for (int t = 0; t <= i + j; ++t) {
        double v = pb * E[i][j - 1][t];
        if (t > 0) v += oo2p * E[i][j - 1][t - 1];
        if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        E[i][j][t] = v;
      }
  }
}

__host__ __device__ constexpr int cart_x(int l, int k) {
  // component k of shell l in rchem's order (shell.rs:1-12): powers (l+1-a, a-b, b-1)
  int idx = 0;
  for (int a = 1; a <= l + 1; ++a)
    for (int b = 1; b <= a; ++b) {
      if (idx == k) return l + 1 - a;
      ++idx;
    }
  return 0;
}
__host__ __device__ constexpr int cart_y(int l, int k) {
  int idx = 0;
  for (int a = 1; a <= l + 1; ++a)
    for (int b = 1; b <= a; ++b) {
      if (idx == k) return a - b;
      ++idx;
    }
  return 0;
}
__host__ __device__ constexpr int cart_z(int l, int k) {
  int idx = 0;
  for (int a = 1; a <= l + 1; ++a)
    for (int b = 1; b <= a; ++b) {
      if (idx == k) return b - 1;
      ++idx;
    }
  return 0;
}

// One thread per shell pair of class (LA, LB): pairs [cls_begin[c], cls_begin[c + 1]),
// c = 3 LA + LB.  Every loop has compile-time bounds, so the Hermite
// coefficients and the R_tuv table live in registers; the Boys block of order
// LA + LB is staged in LDS.
template <int LA, int LB>
__global__ __launch_bounds__(128) void one_electron_kernel(const OneElArgs a) {
  constexpr int L = LA + LB;
  constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2;
  __shared__ double boys_tab[kBoysStageSize];
  stage_boys<L>(boys_tab, a.boys_table);
  __syncthreads();
  const int c = 3 * LA + LB;
  const int pid = a.cls_begin[c] + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (pid >= a.cls_begin[c + 1]) return;
  const int2 pr = a.pairs[pid];
  const ShellDev sa = a.shells[pr.x], sb = a.shells[pr.y];
  double s[NA][NB], t[NA][NB], v[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) s[i][j] = t[i][j] = v[i][j] = 0.0;
  const double AB2 = (sa.xyz[0] - sb.xyz[0]) * (sa.xyz[0] - sb.xyz[0]) +
                     (sa.xyz[1] - sb.xyz[1]) * (sa.xyz[1] - sb.xyz[1]) +
                     (sa.xyz[2] - sb.xyz[2]) * (sa.xyz[2] - sb.xyz[2]);
  int c0 = 0, c1 = a.natom;
  if (a.shell_mol) {  // the pair's own molecule's nuclei
    const int2 r = a.mol_atoms[a.shell_mol[pr.x]];
    c0 = r.x;
    c1 = r.y;
  }
  double E[3][LA + 1][LB + 3][L + 3];
  for (int ip = 0; ip < sa.nprim; ++ip)
    for (int jp = 0; jp < sb.nprim; ++jp) {
      const double al = a.exps[sa.p0 + ip], be = a.exps[sb.p0 + jp];
      const double p = al + be, oop = 1.0 / p;
      const double cc = a.ccoef[sa.p0 + ip] * a.ccoef[sb.p0 + jp];
      const double kab = exp(-al * be * oop * AB2);
      double P[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        P[x] = (al * sa.xyz[x] + be * sb.xyz[x]) * oop;
        hermite_e_t<LA, LB + 2>(P[x] - sa.xyz[x], P[x] - sb.xyz[x], 0.5 * oop, E[x]);
      }
      const double sq = sqrt(3.14159265358979323846 * oop);
      const double pre_s = cc * kab * sq * sq * sq;
      // overlap and kinetic (b shifted by +-2, cints.c:152-177)
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const int A[3] = {cart_x(LA, i), cart_y(LA, i), cart_z(LA, i)};
          const int B[3] = {cart_x(LB, j), cart_y(LB, j), cart_z(LB, j)};
          double ov[3], ovp2[3], ovm2[3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            ov[x] = E[x][A[x]][B[x]][0];
            ovp2[x] = E[x][A[x]][B[x] + 2][0];
            ovm2[x] = B[x] >= 2 ? E[x][A[x]][B[x] >= 2 ? B[x] - 2 : 0][0] : 0.0;
          }
          const double s0 = ov[0] * ov[1] * ov[2];
          s[i][j] += pre_s * s0;
          const int lbt = B[0] + B[1] + B[2];
          const double t1 =
              ovp2[0] * ov[1] * ov[2] + ov[0] * ovp2[1] * ov[2] + ov[0] * ov[1] * ovp2[2];
          const double t2 = B[0] * (B[0] - 1) * ovm2[0] * ov[1] * ov[2] +
                            B[1] * (B[1] - 1) * ov[0] * ovm2[1] * ov[2] +
                            B[2] * (B[2] - 1) * ov[0] * ov[1] * ovm2[2];
          t[i][j] += pre_s * (be * (2 * lbt + 3) * s0 - 2.0 * be * be * t1 - 0.5 * t2);
        }
      // nuclear attraction, one nucleus at a time: R_tuv by the Hermite recursion
      const double pre_v0 = -cc * kab * 2.0 * 3.14159265358979323846 * oop;
      for (int cn = c0; cn < c1; ++cn) {
        const double PC[3] = {P[0] - a.atom_xyz[3 * cn], P[1] - a.atom_xyz[3 * cn + 1],
                              P[2] - a.atom_xyz[3 * cn + 2]};
        const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double F[L + 1];
        boys_exact<L, double>(T, boys_tab, F);
        double R[L + 1][L + 1][L + 1][L + 1];
        double m2p = 1.0;
#pragma unroll
        for (int n = 0; n <= L; ++n) {
          R[n][0][0][0] = m2p * F[n];
          m2p *= -2.0 * p;
        }
#pragma unroll
        for (int n = L - 1; n >= 0; --n) {
#pragma unroll
          for (int tt = 0; tt <= L - n; ++tt)
#pragma unroll
            for (int uu = 0; uu <= L - n - tt; ++uu)
#pragma unroll
              for (int vv = 0; vv <= L - n - tt - uu; ++vv) {
                if (tt + uu + vv == 0) continue;
                double r;
                if (tt > 0) {
                  r = PC[0] * R[n + 1][tt - 1][uu][vv];
                  if (tt > 1) r += (tt - 1) * R[n + 1][tt > 1 ? tt - 2 : 0][uu][vv];
                } else if (uu > 0) {
                  r = PC[1] * R[n + 1][tt][uu - 1][vv];
                  if (uu > 1) r += (uu - 1) * R[n + 1][tt][uu > 1 ? uu - 2 : 0][vv];
                } else {
                  r = PC[2] * R[n + 1][tt][uu][vv - 1];
                  if (vv > 1) r += (vv - 1) * R[n + 1][tt][uu][vv > 1 ? vv - 2 : 0];
                }
                R[n][tt][uu][vv] = r;
              }
        }
        const double pre_v = a.atom_z[cn] * pre_v0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            const int A[3] = {cart_x(LA, i), cart_y(LA, i), cart_z(LA, i)};
            const int B[3] = {cart_x(LB, j), cart_y(LB, j), cart_z(LB, j)};
            double sum = 0.0;
#pragma unroll
            for (int tt = 0; tt <= A[0] + B[0]; ++tt)
#pragma unroll
              for (int uu = 0; uu <= A[1] + B[1]; ++uu)
#pragma unroll
                for (int vv = 0; vv <= A[2] + B[2]; ++vv)
                  sum += E[0][A[0]][B[0]][tt] * E[1][A[1]][B[1]][uu] * E[2][A[2]][B[2]][vv] *
                         R[0][tt][uu][vv];
            v[i][j] += pre_v * sum;
          }
      }
    }
  const size_t n = static_cast<size_t>(a.nbf);
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int ca[3] = {cart_x(LA, i), cart_y(LA, i), cart_z(LA, i)};
      int cb[3] = {cart_x(LB, j), cart_y(LB, j), cart_z(LB, j)};
      const double f = comp_factor(ca) * comp_factor(cb);
      const size_t mu = sa.f0 + i, nu = sb.f0 + j;
      if (a.S) { a.S[mu * n + nu] = f * s[i][j]; a.S[nu * n + mu] = f * s[i][j]; }
      if (a.T) { a.T[mu * n + nu] = f * t[i][j]; a.T[nu * n + mu] = f * t[i][j]; }
      if (a.V) { a.V[mu * n + nu] = f * v[i][j]; a.V[nu * n + mu] = f * v[i][j]; }
    }
}

// ---------------------------------------------------------------------------
// Nuclear gradient of the one-electron energy terms
//   sum_{mu nu} [2 D (dT + dV) - 2 W dS]
// with W the energy-weighted density.  A basis function's derivative with
// respect to its centre is 2 alpha |a + 1_k> - a_k |a - 1_k> (same
// normalisation as the undifferentiated function).  S and T are invariant to
// translating both centres (d/dB = -d/dA); V_C adds the operator term
// d/dC = -(d/dA + d/dB).  One thread per (shell pair a >= b, nucleus C),
// nucleus fastest; the C = 0 thread also adds the S and T terms.
struct OneElGradArgs {
  const ShellDev* shells;
  const double* exps;
  const double* ccoef;
  const int2* pairs;
  int npairs;
  int nbf;
  int natom;
  const double* atom_xyz;
  const double* atom_z;
  const double* boys_table;
  const int* shell_atom;
  const double* D;
  const double* W;
  double* grad;
  int cls_begin[10];  // pairs sorted by class 3 la + lb: class c is [cls_begin[c], cls_begin[c + 1])
};

constexpr int kIg = kLmax1 + 2;      // i range for a + 1
constexpr int kTg = 2 * kLmax1 + 4;  // Hermite t range for (a + 1, b + 2)
constexpr int kRg = 2 * kLmax1 + 2;  // R_tuv index range for L = la + lb + 1

template <int IMAX, int JMAX>
__device__ __forceinline__ void hermite_eg_t(double pa, double pb, double oo2p,
                                             double (&E)[kIg][kE][kTg]) {
  static_assert(IMAX < kIg && JMAX < kE && IMAX + JMAX < kTg, "Hermite table bounds");
#pragma unroll
  for (int i = 0; i <= IMAX; ++i)
#pragma unroll
    for (int j = 0; j <= JMAX; ++j)
#pragma unroll
      for (int t = 0; t < kTg; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
#pragma unroll
  for (int i = 0; i <= IMAX; ++i) {
    if (i > 0) {
#pragma unroll
      for (int t = 0; t <= i; ++t) {
        double v = pa * E[i - 1][0][t];
        if (t > 0) v += oo2p * E[i - 1][0][t > 0 ? t - 1 : 0];
        if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1 < kTg ? t + 1 : 0];
        E[i][0][t] = v;
      }
    }
#pragma unroll
    for (int j = 1; j <= JMAX; ++j)
#pragma unroll
      for (int t = 0; t <= i + j; ++t) {
        double v = pb * E[i][j - 1][t];
        if (t > 0) v += oo2p * E[i][j - 1][t > 0 ? t - 1 : 0];
        if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1 < kTg ? t + 1 : 0];
        E[i][j][t] = v;
      }
  }
}

__device__ __forceinline__ double ovl_g(const double (&E)[3][kIg][kE][kTg], const int* A,
                                        const int* B) {
  return E[0][A[0]][B[0]][0] * E[1][A[1]][B[1]][0] * E[2][A[2]][B[2]][0];
}

// kinetic energy integral over the Hermite overlaps (cints.c:152-177 form)
__device__ __forceinline__ double kin_g(const double (&E)[3][kIg][kE][kTg], const int* A,
                                        const int* B, double be) {
  const double s0 = ovl_g(E, A, B);
  double t1 = 0.0, t2 = 0.0;
  for (int x = 0; x < 3; ++x) {
    int Bp[3] = {B[0], B[1], B[2]};
    Bp[x] += 2;
    t1 += ovl_g(E, A, Bp);
    if (B[x] >= 2) {
      int Bm[3] = {B[0], B[1], B[2]};
      Bm[x] -= 2;
      t2 += B[x] * (B[x] - 1) * ovl_g(E, A, Bm);
    }
  }
  const int lb = B[0] + B[1] + B[2];
  return be * (2 * lb + 3) * s0 - 2.0 * be * be * t1 - 0.5 * t2;
}

__device__ __forceinline__ double nuc_g(const double (&E)[3][kIg][kE][kTg],
                                        const double (&R)[kRg][kRg][kRg], const int* A,
                                        const int* B) {
  double sum = 0.0;
  for (int t = 0; t <= A[0] + B[0]; ++t)
    for (int u = 0; u <= A[1] + B[1]; ++u)
      for (int v = 0; v <= A[2] + B[2]; ++v)
        sum += E[0][A[0]][B[0]][t] * E[1][A[1]][B[1]][u] * E[2][A[2]][B[2]][v] * R[t][u][v];
  return sum;
}

// One thread per (shell pair of class (LA, LB), nucleus), nucleus fastest.
// The class is a template argument so every loop bound is known at compile
// time and the Hermite and R_tuv arrays stay in registers; the Boys block of
// order LA + LB + 1 is staged in LDS.
template <int LA, int LB>
__device__ __forceinline__ void one_electron_grad_terms(const OneElGradArgs& a, int pid, int c,
                                                        const double* __restrict__ boys_tab,
                                                        double* gA, double* gB, double* gC) {
  constexpr int L = LA + LB + 1;
  constexpr int na = (LA + 1) * (LA + 2) / 2, nb = (LB + 1) * (LB + 2) / 2;
  const int2 pr = a.pairs[pid];
  const ShellDev sa = a.shells[pr.x], sb = a.shells[pr.y];
  const double f = pr.x != pr.y ? 2.0 : 1.0;  // the (b, a) block of the sum
  const double AB2 = (sa.xyz[0] - sb.xyz[0]) * (sa.xyz[0] - sb.xyz[0]) +
                     (sa.xyz[1] - sb.xyz[1]) * (sa.xyz[1] - sb.xyz[1]) +
                     (sa.xyz[2] - sb.xyz[2]) * (sa.xyz[2] - sb.xyz[2]);
  const double Cx = a.atom_xyz[3 * c], Cy = a.atom_xyz[3 * c + 1], Cz = a.atom_xyz[3 * c + 2];
  const size_t n = static_cast<size_t>(a.nbf);
  double E[3][kIg][kE][kTg];
  double R[kRg][kRg][kRg], Rn[kRg][kRg][kRg];
  double F[L + 1];
  for (int ip = 0; ip < sa.nprim; ++ip)
    for (int jp = 0; jp < sb.nprim; ++jp) {
      const double al = a.exps[sa.p0 + ip], be = a.exps[sb.p0 + jp];
      const double p = al + be, oop = 1.0 / p;
      const double cc = a.ccoef[sa.p0 + ip] * a.ccoef[sb.p0 + jp];
      const double kab = exp(-al * be * oop * AB2);
      double P[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        P[x] = (al * sa.xyz[x] + be * sb.xyz[x]) * oop;
        hermite_eg_t<LA + 1, LB + 2>(P[x] - sa.xyz[x], P[x] - sb.xyz[x], 0.5 * oop, E[x]);
      }
      // Hermite Coulomb integrals R_tuv(p, P - C), level by level (R_{n} from R_{n+1})
      const double PC[3] = {P[0] - Cx, P[1] - Cy, P[2] - Cz};
      const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
      boys_exact<L, double>(T, boys_tab, F);
      double m2p = 1.0;
#pragma unroll
      for (int k = 0; k < L; ++k) m2p *= -2.0 * p;
      Rn[0][0][0] = m2p * F[L];
#pragma unroll
      for (int lev = L - 1; lev >= 0; --lev) {
        m2p /= -2.0 * p;
#pragma unroll
        for (int tt = 0; tt <= L - lev; ++tt)
#pragma unroll
          for (int uu = 0; uu <= L - lev - tt; ++uu)
#pragma unroll
            for (int vv = 0; vv <= L - lev - tt - uu; ++vv) {
              double r;
              if (tt + uu + vv == 0) {
                r = m2p * F[lev];
              } else if (tt > 0) {
                r = PC[0] * Rn[tt - 1][uu][vv];
                if (tt > 1) r += (tt - 1) * Rn[tt > 1 ? tt - 2 : 0][uu][vv];
              } else if (uu > 0) {
                r = PC[1] * Rn[tt][uu - 1][vv];
                if (uu > 1) r += (uu - 1) * Rn[tt][uu > 1 ? uu - 2 : 0][vv];
              } else {
                r = PC[2] * Rn[tt][uu][vv - 1];
                if (vv > 1) r += (vv - 1) * Rn[tt][uu][vv > 1 ? vv - 2 : 0];
              }
              R[tt][uu][vv] = r;
            }
#pragma unroll
        for (int tt = 0; tt <= L - lev; ++tt)
#pragma unroll
          for (int uu = 0; uu <= L - lev - tt; ++uu)
#pragma unroll
            for (int vv = 0; vv <= L - lev - tt - uu; ++vv) Rn[tt][uu][vv] = R[tt][uu][vv];
      }
      const double sq = sqrt(3.14159265358979323846 * oop);
      const double pre_s = cc * kab * sq * sq * sq;
      const double pre_v = -a.atom_z[c] * cc * kab * 2.0 * 3.14159265358979323846 * oop;
#pragma unroll
      for (int i = 0; i < na; ++i)
#pragma unroll
        for (int j = 0; j < nb; ++j) {
          const int A[3] = {cart_x(LA, i), cart_y(LA, i), cart_z(LA, i)};
          const int B[3] = {cart_x(LB, j), cart_y(LB, j), cart_z(LB, j)};
          const double w = f * comp_factor(A) * comp_factor(B);
          const size_t mu = sa.f0 + i, nu = sb.f0 + j;
          const double d2 = 2.0 * a.D[mu * n + nu], w2 = 2.0 * a.W[mu * n + nu];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            int Ap[3] = {A[0], A[1], A[2]}, Am[3] = {A[0], A[1], A[2]};
            int Bp[3] = {B[0], B[1], B[2]}, Bm[3] = {B[0], B[1], B[2]};
            Ap[k] += 1;
            Am[k] = Am[k] > 0 ? Am[k] - 1 : 0;
            Bp[k] += 1;
            Bm[k] = Bm[k] > 0 ? Bm[k] - 1 : 0;
            double dVA = 2.0 * al * nuc_g(E, R, Ap, B);
            if (A[k] > 0) dVA -= A[k] * nuc_g(E, R, Am, B);
            double dVB = 2.0 * be * nuc_g(E, R, A, Bp);
            if (B[k] > 0) dVB -= B[k] * nuc_g(E, R, A, Bm);
            const double vA = w * d2 * pre_v * dVA, vB = w * d2 * pre_v * dVB;
            gA[k] += vA;
            gB[k] += vB;
            gC[k] -= vA + vB;
            if (c == 0) {
              double dS = 2.0 * al * ovl_g(E, Ap, B);
              double dT = 2.0 * al * kin_g(E, Ap, B, be);
              if (A[k] > 0) {
                dS -= A[k] * ovl_g(E, Am, B);
                dT -= A[k] * kin_g(E, Am, B, be);
              }
              const double st = w * pre_s * (d2 * dT - w2 * dS);
              gA[k] += st;
              gB[k] -= st;
            }
          }
        }
    }
}

__device__ __forceinline__ double wave_sum_1(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Lanes are (pair, nucleus) with the nucleus fastest, so a wave mostly holds
// one pair: its d/dA and d/dB terms are summed across the wave first (one
// atomic per component instead of 64 on the same few addresses).
template <int LA, int LB>
__global__ __launch_bounds__(128) void one_electron_grad_kernel(const OneElGradArgs a) {
  __shared__ double boys_tab[kBoysStageSize];
  stage_boys<LA + LB + 1>(boys_tab, a.boys_table);
  __syncthreads();
  const int c_cls = 3 * LA + LB;
  const int p0 = a.cls_begin[c_cls], np = a.cls_begin[c_cls + 1] - p0;
  const long tid = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = tid < static_cast<long>(np) * a.natom;
  const int pid = valid ? p0 + static_cast<int>(tid / a.natom) : p0;
  const int c = valid ? static_cast<int>(tid % a.natom) : 0;
  double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
  if (valid) one_electron_grad_terms<LA, LB>(a, pid, c, boys_tab, gA, gB, gC);
  const int2 pr = a.pairs[pid];
  const int atA = a.shell_atom[pr.x], atB = a.shell_atom[pr.y];
  const int pid0 = __shfl(pid, 0, 64);
  const bool uniform = __all(!valid || pid == pid0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (uniform) {
      const double sA = wave_sum_1(gA[k]), sB = wave_sum_1(gB[k]);
      if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(&a.grad[3 * atA + k], sA);
        unsafeAtomicAdd(&a.grad[3 * atB + k], sB);
      }
    } else if (valid) {
      unsafeAtomicAdd(&a.grad[3 * atA + k], gA[k]);
      unsafeAtomicAdd(&a.grad[3 * atB + k], gB[k]);
    }
    if (valid) unsafeAtomicAdd(&a.grad[3 * c + k], gC[k]);
  }
}

template <int LA, int LB>
hipError_t one_electron_grad_class(const OneElGradArgs& a, hipStream_t s) {
  const int c = 3 * LA + LB;
  const long total = static_cast<long>(a.cls_begin[c + 1] - a.cls_begin[c]) * a.natom;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL((one_electron_grad_kernel<LA, LB>),
                     dim3(static_cast<unsigned>((total + 127) / 128)), dim3(128), 0, s, a);
  return hipGetLastError();
}

// a.pairs must be sorted by class 3 la + lb with a.cls_begin its boundaries.
hipError_t one_electron_grad_launch(const OneElGradArgs& a, hipStream_t s) {
  if (a.npairs <= 0 || a.natom <= 0) return hipSuccess;
  hipError_t e;
  if ((e = one_electron_grad_class<0, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<0, 1>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<0, 2>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<1, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<1, 1>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<1, 2>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<2, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_grad_class<2, 1>(a, s)) != hipSuccess) return e;
  return one_electron_grad_class<2, 2>(a, s);
}

template <int LA, int LB>
hipError_t one_electron_class(const OneElArgs& a, hipStream_t s) {
  const int c = 3 * LA + LB;
  const int cnt = a.cls_begin[c + 1] - a.cls_begin[c];
  if (cnt <= 0) return hipSuccess;
  hipLaunchKernelGGL((one_electron_kernel<LA, LB>), dim3((cnt + 127) / 128), dim3(128), 0, s, a);
  return hipGetLastError();
}

// a.pairs must be sorted by class 3 la + lb with a.cls_begin its boundaries.
hipError_t one_electron_launch(const OneElArgs& a, hipStream_t s) {
  if (a.npairs <= 0) return hipSuccess;
  hipError_t e;
  if ((e = one_electron_class<0, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<0, 1>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<0, 2>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<1, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<1, 1>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<1, 2>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<2, 0>(a, s)) != hipSuccess) return e;
  if ((e = one_electron_class<2, 1>(a, s)) != hipSuccess) return e;
  return one_electron_class<2, 2>(a, s);
}

}  // namespace rchem

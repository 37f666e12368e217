// ERI launch tables and the small digestion kernels that follow the ERIs.
// The per-class kernels live in one translation unit per (la lb|lc ld) class
// (generated/eri_c_*.hip, written by gen_eri.py) so they compile in parallel
// and instantiate only the variants each class needs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "types.hpp"

namespace rchem {

using EriLaunchFn = hipError_t (*)(const EriArgs&, int, hipStream_t);

#include "generated/eri_dispatch.inc"

// Jh[fa + ia, fb + ib] = Jp of pair p (each shell pair appears once, so the
// writes do not collide); one thread per (pair, component).
__global__ void jp_scatter_kernel(const PairRec* __restrict__ pairs, const double* __restrict__ Jp,
                                  const JpClass* __restrict__ cls, const int* __restrict__ ncomp,
                                  int ncls, int npairs, double* __restrict__ Jh, int nbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int X = 0;
  while (X + 1 < ncls && cls[X + 1].begin <= p) ++X;
  const JpClass jc = cls[X];
  const int nb = ncomp[2 * X + 1];
  const int nc = ncomp[2 * X] * nb;
  const PairRec pr = pairs[p];
  const size_t n = static_cast<size_t>(nbf);
  for (int k = 0; k < nc; ++k) {
    const double v = Jp[jc.base + k * jc.n + (p - jc.begin)];
    Jh[(pr.fa + k / nb) * n + pr.fb + k % nb] += v;
  }
}

// Dp of pair p = D[fa + ia, fb + ib], in the layout of Jp: the ket pairs of a
// tile are consecutive, so a wave reads its 64 ket density blocks coalesced
// instead of gathering them from all over D.  One thread per pair.
__global__ void dpair_gather_kernel(const PairRec* __restrict__ pairs, const double* __restrict__ D,
                                    const JpClass* __restrict__ cls, const int* __restrict__ ncomp,
                                    int ncls, int npairs, double* __restrict__ Dp, int nbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int X = 0;
  while (X + 1 < ncls && cls[X + 1].begin <= p) ++X;
  const JpClass jc = cls[X];
  const int nb = ncomp[2 * X + 1];
  const int nc = ncomp[2 * X] * nb;
  const PairRec pr = pairs[p];
  const size_t n = static_cast<size_t>(nbf);
  for (int k = 0; k < nc; ++k)
    Dp[jc.base + k * jc.n + (p - jc.begin)] = D[(pr.fa + k / nb) * n + pr.fb + k % nb];
}

// J = Jh + Jh^T, K = Kh + Kh^T (the lower-half accumulators of eri_kernel).
__global__ void symmetrize_kernel(const double* __restrict__ Jh, const double* __restrict__ Kh,
                                  double* __restrict__ J, double* __restrict__ K, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const size_t ij = static_cast<size_t>(i) * n + j, ji = static_cast<size_t>(j) * n + i;
  J[ij] = Jh[ij] + Jh[ji];
  K[ij] = Kh[ij] + Kh[ji];
}

// Block-diagonal versions for a multi-molecule basis (quartets within a
// molecule only, so J and K are block-diagonal): blockIdx.y = molecule m,
// elements of its nb[m] x nb[m] block at rows and columns f0[m] ...
__global__ void symmetrize_blocks_kernel(const double* __restrict__ Jh,
                                         const double* __restrict__ Kh, double* __restrict__ J,
                                         double* __restrict__ K, int n, const int* __restrict__ f0,
                                         const int* __restrict__ nb) {
  const int m = blockIdx.y, b = nb[m], o = f0[m];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < b * b; k += gridDim.x * blockDim.x) {
    const size_t i = o + k / b, j = o + k % b;
    const size_t ij = i * n + j, ji = j * n + i;
    J[ij] = Jh[ij] + Jh[ji];
    K[ij] = Kh[ij] + Kh[ji];
  }
}

__global__ void fill_blocks_kernel(double* __restrict__ A, double* __restrict__ B, int n,
                                   const int* __restrict__ f0, const int* __restrict__ nb) {
  const int m = blockIdx.y, b = nb[m], o = f0[m];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < b * b; k += gridDim.x * blockDim.x) {
    const size_t ij = static_cast<size_t>(o + k / b) * n + o + k % b;
    A[ij] = 0.0;
    B[ij] = 0.0;
  }
}

hipError_t jp_scatter_launch(const PairRec* pairs, const double* Jp, const JpClass* cls,
                             const int* ncomp, int ncls, int npairs, double* Jh, int nbf,
                             hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(jp_scatter_kernel, dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, Jp,
                     cls, ncomp, ncls, npairs, Jh, nbf);
  return hipGetLastError();
}

hipError_t dpair_gather_launch(const PairRec* pairs, const double* D, const JpClass* cls,
                               const int* ncomp, int ncls, int npairs, double* Dp, int nbf,
                               hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(dpair_gather_kernel, dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, D,
                     cls, ncomp, ncls, npairs, Dp, nbf);
  return hipGetLastError();
}

hipError_t symmetrize_launch(const double* Jh, const double* Kh, double* J, double* K, int n,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  dim3 grid((n + 255) / 256, n);
  hipLaunchKernelGGL(symmetrize_kernel, grid, dim3(256), 0, s, Jh, Kh, J, K, n);
  return hipGetLastError();
}

// J = Jh + Jh^T, K = Kh + Kh^T on the diagonal blocks only (nmol molecules,
// at most nmax functions each); the off-diagonal blocks of J and K are not
// written.
hipError_t symmetrize_blocks_launch(const double* Jh, const double* Kh, double* J, double* K,
                                    int n, const int* f0, const int* nb, int nmol, int nmax,
                                    hipStream_t s) {
  if (nmol <= 0) return hipSuccess;
  const unsigned gx = static_cast<unsigned>(std::min(64, (nmax * nmax + 255) / 256));
  hipLaunchKernelGGL(symmetrize_blocks_kernel, dim3(gx, nmol), dim3(256), 0, s, Jh, Kh, J, K, n,
                     f0, nb);
  return hipGetLastError();
}

// Clears the diagonal blocks of A and B.
hipError_t fill_blocks_launch(double* A, double* B, int n, const int* f0, const int* nb, int nmol,
                              int nmax, hipStream_t s) {
  if (nmol <= 0) return hipSuccess;
  const unsigned gx = static_cast<unsigned>(std::min(64, (nmax * nmax + 255) / 256));
  hipLaunchKernelGGL(fill_blocks_kernel, dim3(gx, nmol), dim3(256), 0, s, A, B, n, f0, nb);
  return hipGetLastError();
}

}  // namespace rchem

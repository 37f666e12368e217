// Small elementwise and reduction kernels of the SCF step (gfx950).
//   fill:          graph-safe clearing of accumulators (no hipMemsetAsync);
//   fock:          F = H + 2J - K (rchem.rs:110-127);
//   energy:        E = sum D o (H + F) (rchem.rs:106-108), one workgroup,
//                  fixed reduction order, so it is deterministic;
//   scale_columns: U diag(lambda^-1/2), the middle of Loewdin X (rchem.rs:44-51);
//   antisym:       E = A - A^T, the DIIS commutator FDS - SDF from A = FDS;
//   purify_*:      canonical purification (Palser & Manolopoulos, Phys. Rev. B
//                  58, 12704 (1998)) of an orthonormal-basis Fock matrix F into
//                  the projector P onto its nocc lowest eigenvectors, the
//                  density rchem forms from eigh (rchem.rs:74-78, 101-104),
//                  with GEMMs (MFMA) in place of the eigensolver:
//                    gersh:  Gershgorin bounds per row, lo_i/hi_i = F_ii -/+ sum_{j!=i} |F_ij|;
//                    params: e_min = min lo, e_max = max hi, mu = tr F / n,
//                            lambda = min(N / (e_max - mu), (n - N) / (mu - e_min));
//                    start:  P_0 = lambda / n (mu I - F) + N / n I (spectrum in [0, 1], trace N);
//                    coef:   c = tr(P^2 - P^3) / tr(P - P^2) (and tr(P - P^2) for the host);
//                    step:   c <= 1/2: P <- ((1 - 2c) P + (1 + c) P^2 - P^3) / (1 - c),
//                            c >  1/2: P <- ((1 + c) P^2 - P^3) / c,
//                  maps that are monotonic on [0, 1], fix 0 and 1, keep tr P = N,
//                  and drive every eigenvalue to 0 or 1.
#include <hip/hip_runtime.h>

namespace rchem {

__global__ void fill_kernel(double* __restrict__ p, size_t n, double v) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void fock_kernel(const double* __restrict__ H, const double* __restrict__ J,
                            const double* __restrict__ K, double* __restrict__ F, size_t n2) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n2;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    F[i] = H[i] + 2.0 * J[i] - K[i];
}

__global__ __launch_bounds__(1024) void energy_kernel(const double* __restrict__ D,
                                                      const double* __restrict__ H,
                                                      const double* __restrict__ F,
                                                      double* __restrict__ out, size_t n2) {
  __shared__ double part[1024];
  double s = 0.0;
  for (size_t i = threadIdx.x; i < n2; i += blockDim.x) s += D[i] * (H[i] + F[i]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

__global__ void scale_columns_kernel(const double* __restrict__ U, const double* __restrict__ w,
                                     double* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const size_t k = static_cast<size_t>(j) * n + i;  // column-major (i, j)
  out[k] = U[k] / sqrt(w[j]);
}

__global__ void antisym_kernel(const double* __restrict__ A, double* __restrict__ E, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  E[static_cast<size_t>(j) * n + i] = A[static_cast<size_t>(j) * n + i] - A[static_cast<size_t>(i) * n + j];
}

__device__ __forceinline__ double wave_sum64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// one wave per row, four rows per workgroup (F symmetric)
__global__ __launch_bounds__(256) void gersh_kernel(const double* __restrict__ F, int n,
                                                    double* __restrict__ lo,
                                                    double* __restrict__ hi) {
  const int row = blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;  // wave-uniform
  const double* f = F + static_cast<size_t>(row) * n;
  double r = 0.0;
  for (int j = lane; j < n; j += 64) r += j == row ? 0.0 : fabs(f[j]);
  r = wave_sum64(r);
  if (lane == 0) {
    lo[row] = f[row] - r;
    hi[row] = f[row] + r;
  }
}

// par = {lambda / n, mu, N / n}
__global__ __launch_bounds__(1024) void purify_params_kernel(const double* __restrict__ F,
                                                             const double* __restrict__ lo,
                                                             const double* __restrict__ hi, int n,
                                                             int nocc, double* __restrict__ par) {
  __shared__ double smin[1024], smax[1024], ssum[1024];
  double mn = HUGE_VAL, mx = -HUGE_VAL, tr = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    mn = fmin(mn, lo[i]);
    mx = fmax(mx, hi[i]);
    tr += F[static_cast<size_t>(i) * n + i];
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = tr;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mu = ssum[0] / n;
    const double up = smax[0] - mu, down = mu - smin[0];
    double lam = HUGE_VAL;
    if (up > 0.0) lam = fmin(lam, nocc / up);
    if (down > 0.0) lam = fmin(lam, (n - nocc) / down);
    if (!(lam < HUGE_VAL)) lam = 0.0;  // F = mu I: P_0 = N / n I
    par[0] = lam / n;
    par[1] = mu;
    par[2] = static_cast<double>(nocc) / n;
  }
}

__global__ void purify_start_kernel(const double* __restrict__ F, const double* __restrict__ par,
                                    int n, double* __restrict__ P) {
  const size_t n2 = static_cast<size_t>(n) * n;
  const double a = par[0], mu = par[1], diag = par[2];
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const bool d = (k / n) == (k % n);
    P[k] = a * ((d ? mu : 0.0) - F[k]) + (d ? diag : 0.0);
  }
}

// coef = {c, tr(P - P^2)}; c = 1/2 (McWeeny's step) once tr(P - P^2) is at
// rounding level, where the ratio carries no information
__global__ __launch_bounds__(1024) void purify_coef_kernel(const double* __restrict__ P,
                                                           const double* __restrict__ P2,
                                                           const double* __restrict__ P3, int n,
                                                           double* __restrict__ coef) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const size_t d = static_cast<size_t>(i) * n + i;
    a += P[d] - P2[d];
    b += P2[d] - P3[d];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double t1 = s1[0], t2 = s2[0];
    double c = 0.5;
    if (t1 > 1e-13 * n) c = fmin(fmax(t2 / t1, 1e-3), 1.0 - 1e-3);
    coef[0] = c;
    coef[1] = t1;
  }
}

__global__ void purify_step_kernel(double* __restrict__ P, const double* __restrict__ P2,
                                   const double* __restrict__ P3, const double* __restrict__ coef,
                                   size_t n2) {
  const double c = coef[0];
  double a, b, d;  // P <- a P + b P^2 + d P^3
  if (c <= 0.5) {
    a = (1.0 - 2.0 * c) / (1.0 - c);
    b = (1.0 + c) / (1.0 - c);
    d = -1.0 / (1.0 - c);
  } else {
    a = 0.0;
    b = (1.0 + c) / c;
    d = -1.0 / c;
  }
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    P[k] = a * P[k] + b * P2[k] + d * P3[k];
}

static unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b == 0) b = 1;
  return static_cast<unsigned>(b);
}

hipError_t fill_launch(double* p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t fock_launch(const double* H, const double* J, const double* K, double* F, size_t n2,
                       hipStream_t s) {
  hipLaunchKernelGGL(fock_kernel, dim3(grid_for(n2)), dim3(256), 0, s, H, J, K, F, n2);
  return hipGetLastError();
}

hipError_t energy_launch(const double* D, const double* H, const double* F, double* out,
                         size_t n2, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, s, D, H, F, out, n2);
  return hipGetLastError();
}

hipError_t scale_columns_launch(const double* U, const double* w, double* out, int n,
                                hipStream_t s) {
  hipLaunchKernelGGL(scale_columns_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, U, w, out,
                     n);
  return hipGetLastError();
}

}  // namespace rchem

namespace rchem {
hipError_t antisym_launch(const double* A, double* E, int n, hipStream_t s) {
  hipLaunchKernelGGL(antisym_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, A, E, n);
  return hipGetLastError();
}
}  // namespace rchem

namespace rchem {
// Start of the purification of F (n x n, symmetric) for nocc occupied
// orbitals: P = P_0.  lo, hi: n doubles of scratch; par: 3 doubles.
hipError_t purify_start_launch(const double* F, int n, int nocc, double* lo, double* hi,
                               double* par, double* P, hipStream_t s) {
  hipLaunchKernelGGL(gersh_kernel, dim3((n + 3) / 4), dim3(256), 0, s, F, n, lo, hi);
  hipLaunchKernelGGL(purify_params_kernel, dim3(1), dim3(1024), 0, s, F, lo, hi, n, nocc, par);
  hipLaunchKernelGGL(purify_start_kernel, dim3(grid_for(static_cast<size_t>(n) * n)), dim3(256),
                     0, s, F, par, n, P);
  return hipGetLastError();
}

// One purification step given P2 = P P and P3 = P2 P; coef receives
// {c, tr(P - P^2)} of the P before the step.
hipError_t purify_step_launch(double* P, const double* P2, const double* P3, int n,
                              double* coef, hipStream_t s) {
  hipLaunchKernelGGL(purify_coef_kernel, dim3(1), dim3(1024), 0, s, P, P2, P3, n, coef);
  const size_t n2 = static_cast<size_t>(n) * n;
  hipLaunchKernelGGL(purify_step_kernel, dim3(grid_for(n2)), dim3(256), 0, s, P, P2, P3, coef,
                     n2);
  return hipGetLastError();
}
}  // namespace rchem

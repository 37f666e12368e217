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
//   diis_*:        Pulay DIIS on the device (state in ScfDevState), so that an
//                  SCF iteration -- Fock build, energy, commutator, DIIS,
//                  extrapolation, purification, density -- has no host step
//                  and replays as one graph:
//                    push:   the commutator's dot products with the stored ones
//                            (fixed grid, fixed summation order), its max |.|,
//                            and F and the commutator into the ring slot;
//                    solve:  one workgroup: the convergence / precision-switch
//                            decision and [B 1; 1 0] (c, -lambda) = (0, 1) by
//                            Gauss-Jordan elimination with partial pivoting;
//                    extrapolate: F_x = sum_i c_i F_i;
//                    commit: D <- D_new, or (1 - a) D_new + a D (damping).
#include <hip/hip_runtime.h>

#include "types.hpp"

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

namespace rchem {

// sum over a 256-thread workgroup in a fixed order; sh: 4 doubles of LDS
__device__ double block_sum256(double v, double* sh) {
  v = wave_sum64(v);
  __syncthreads();  // the previous use of sh is finished
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__device__ double block_max256(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// part[block * (kDiisMax + 1) + i] = this block's share of <E, E_i> (E_slot = E),
// part[block * (kDiisMax + 1) + kDiisMax] = its max |E|
__global__ __launch_bounds__(256) void diis_push_kernel(const double* __restrict__ E,
                                                        const double* __restrict__ F,
                                                        double* __restrict__ Es,
                                                        double* __restrict__ Fs, size_t n2,
                                                        const ScfDevState* __restrict__ st, int k,
                                                        double* __restrict__ part) {
  __shared__ double sh[4];
  const int count = st->count;
  const int slot = count % k;
  const int m = min(count + 1, k);
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t t0 = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  double* pb = part + static_cast<size_t>(blockIdx.x) * (kDiisMax + 1);
  for (int i = 0; i < m; ++i) {
    // the slot being replaced is never read: its new content is E itself
    const double* Ei = i == slot ? E : Es + static_cast<size_t>(i) * n2;
    double s = 0.0;
    for (size_t e = t0; e < n2; e += stride) s += E[e] * Ei[e];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) pb[i] = s;
  }
  double* Eslot = Es + static_cast<size_t>(slot) * n2;
  double* Fslot = Fs + static_cast<size_t>(slot) * n2;
  double mx = 0.0;
  for (size_t e = t0; e < n2; e += stride) {
    const double x = E[e];
    mx = fmax(mx, fabs(x));
    Eslot[e] = x;
    Fslot[e] = F[e];
  }
  mx = block_max256(mx, sh);
  if (threadIdx.x == 0) pb[kDiisMax] = mx;
}

__global__ __launch_bounds__(64) void diis_solve_kernel(ScfDevState* __restrict__ st,
                                                        const double* __restrict__ part, int nblk,
                                                        const double* __restrict__ energy,
                                                        ScfDevParams p) {
  constexpr int W = kDiisMax + 2;  // augmented row: kDiisMax + 1 unknowns, right-hand side
  __shared__ double A[(kDiisMax + 1) * W];
  __shared__ double red[kDiisMax + 1];
  __shared__ int s_piv, s_bad;
  const int t = threadIdx.x;
  const int k = p.k;
  int count = st->count;
  const int slot = count % k;
  ++count;
  const int m = min(count, k);
  if (t < m || t == kDiisMax) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const double v = part[static_cast<size_t>(b) * (kDiisMax + 1) + t];
      s = t == kDiisMax ? fmax(s, v) : s + v;
    }
    red[t] = s;
  }
  __syncthreads();
  if (t < m) {
    st->B[slot * kDiisMax + t] = red[t];
    st->B[t * kDiisMax + slot] = red[t];
  }
  __syncthreads();
  const int it = st->it;
  const double e_cur = energy[0];
  const double de = e_cur - st->e_prev;
  const double errmax = red[kDiisMax];
  const bool sw = it > 0 && p.mixed && fabs(de) < p.switch_de;
  const bool conv = !sw && it > 0 && p.fp64 && fabs(de) < p.thresh_e && errmax < p.errtol;
  bool singular = false;
  if (!sw && !conv) {
    // [B 1; 1 0] (c, -lambda) = (0, 1)
    const int sz = m + 1;
    for (int r = t; r < sz; r += blockDim.x) {
      for (int j = 0; j < m; ++j)
        A[r * W + j] = r < m ? st->B[r * kDiisMax + j] : 1.0;
      A[r * W + m] = r < m ? 1.0 : 0.0;
      A[r * W + sz] = r < m ? 0.0 : 1.0;
    }
    if (t == 0) s_bad = 0;
    __syncthreads();
    for (int c = 0; c < sz; ++c) {
      if (t == 0) {
        int piv = c;
        for (int r = c + 1; r < sz; ++r)
          if (fabs(A[r * W + c]) > fabs(A[piv * W + c])) piv = r;
        s_piv = piv;
        if (fabs(A[piv * W + c]) < 1e-300) s_bad = 1;
      }
      __syncthreads();
      if (s_bad) break;  // uniform: every thread read the same flag after the barrier
      const int piv = s_piv;
      if (piv != c)
        for (int j = t; j <= sz; j += blockDim.x) {
          const double a = A[c * W + j];
          A[c * W + j] = A[piv * W + j];
          A[piv * W + j] = a;
        }
      __syncthreads();
      for (int r = t; r < sz; r += blockDim.x) {
        if (r == c) continue;
        const double f = A[r * W + c] / A[c * W + c];
        for (int j = c; j <= sz; ++j) A[r * W + j] -= f * A[c * W + j];
      }
      __syncthreads();
    }
    singular = s_bad != 0;
    if (!singular && t < m) st->c[t] = A[t * W + sz] / A[t * W + t];
  }
  __syncthreads();
  if (t < kDiisMax && (sw || conv || singular)) st->c[t] = t == slot ? 1.0 : 0.0;  // F_x = F
  if (t == 0) {
    st->e_cur = e_cur;
    st->de = de;
    st->errmax = errmax;
    st->converged = conv;
    st->switched = sw;
    st->singular = singular;
    st->purify_failed = 0;
    st->it = it + 1;
    st->count = singular ? 0 : count;
    st->m = m;
    st->slot = slot;
    st->damp = !sw && !conv && p.damping > 0.0 && errmax >= p.damping_until;
    st->e_prev = e_cur;
  }
}

__global__ void diis_extrapolate_kernel(const double* __restrict__ Fs, size_t n2,
                                        const ScfDevState* __restrict__ st,
                                        double* __restrict__ Fx) {
  const int m = st->m;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n2;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += st->c[i] * Fs[static_cast<size_t>(i) * n2 + e];
    Fx[e] = s;
  }
}

// D <- D_new (damped: (1 - a) D_new + a D) unless the precision switch keeps D
// or the purification has not converged.  coef holds {c, tr(P - P^2)} of each
// of the purification's `steps` steps (steps = 0: the density came from
// elsewhere, commit); it has converged when tr(P - P^2) fell below tol at
// least two steps before its end, and purify_steps records the first such
// step + 3.  Otherwise purify_failed is raised and the host takes over.
__global__ void scf_commit_kernel(double* __restrict__ D, const double* __restrict__ Dn, size_t n2,
                                  ScfDevState* __restrict__ st, const double* __restrict__ coef,
                                  int steps, double tol, double damping) {
  if (st->switched) return;
  int first = steps;
  for (int s = 0; s < steps; ++s)
    if (coef[2 * s + 1] < tol) {  // false for NaN
      first = s;
      break;
    }
  if (steps > 0 && first + 3 > steps) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->purify_failed = 1;
      st->purify_steps = steps + 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->purify_steps = steps > 0 ? first + 3 : 0;
  const double a = st->damp ? damping : 0.0;
  const double keep = 1.0 - a;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n2;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    D[e] = a > 0.0 ? keep * Dn[e] + a * D[e] : Dn[e];
}

hipError_t diis_push_launch(const double* E, const double* F, double* Es, double* Fs, size_t n2,
                            const ScfDevState* st, int k, double* part, hipStream_t s) {
  hipLaunchKernelGGL(diis_push_kernel, dim3(kDiisBlocks), dim3(256), 0, s, E, F, Es, Fs, n2, st, k,
                     part);
  return hipGetLastError();
}

hipError_t diis_solve_launch(ScfDevState* st, const double* part, const double* energy,
                             const ScfDevParams& p, hipStream_t s) {
  hipLaunchKernelGGL(diis_solve_kernel, dim3(1), dim3(64), 0, s, st, part, kDiisBlocks, energy, p);
  return hipGetLastError();
}

hipError_t diis_extrapolate_launch(const double* Fs, size_t n2, const ScfDevState* st, double* Fx,
                                   hipStream_t s) {
  hipLaunchKernelGGL(diis_extrapolate_kernel, dim3(grid_for(n2)), dim3(256), 0, s, Fs, n2, st, Fx);
  return hipGetLastError();
}

hipError_t scf_commit_launch(double* D, const double* Dn, size_t n2, ScfDevState* st,
                             const double* coef, int steps, double tol, double damping,
                             hipStream_t s) {
  hipLaunchKernelGGL(scf_commit_kernel, dim3(grid_for(n2)), dim3(256), 0, s, D, Dn, n2, st, coef,
                     steps, tol, damping);
  return hipGetLastError();
}

}  // namespace rchem

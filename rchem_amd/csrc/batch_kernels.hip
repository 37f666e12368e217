// Kernels of the batched SCF step: many molecules, each a diagonal block
// [f0_b, f0_b + n_b) of the dense matrices of a multi-molecule basis, and a
// padded copy of that block (nmax x nmax, column-major for rocBLAS/rocSOLVER's
// strided-batched routines; the blocks are symmetric except where noted).
//   pack:        out_b = block_b of A, padding = pad * identity;
//   unpack:      block_b of A = in_b (the off-diagonal blocks stay untouched);
//   mask_cols:   C_b(:, j) = 0 for j >= nocc_b (so D_b = C_b C_b^T over nmax);
//   inv_sqrt_cols: U_b diag(lambda_b)^-1/2 (Loewdin X, padding lambda = 1);
//   energy:      E_b = sum over block b of D (H + F), one workgroup per molecule;
//   antisym:     E_b = A_b - A_b^T;
//   dots:        <e_b, f_b> for a list of pairs of padded blocks, one workgroup each;
//   maxabs:      max |e_b| per molecule;
//   combine:     out_b = sum_i c[b][i] F_i,b (DIIS extrapolation);
//   blend:       block_b of D = (1 - a_b) block_b of D + a_b Dold_b (padded old copy);
//   gather:      out_s = top-left nout x nout of in_idx[s] (drops retired molecules
//                and shrinks the padding);
//   fock:        F = H + 2J - K on the diagonal blocks only.
#include <hip/hip_runtime.h>

namespace rchem {

namespace {

__global__ void pack_kernel(const double* __restrict__ A, size_t n, const int* __restrict__ f0,
                            const int* __restrict__ nb, int nmax, double pad,
                            double* __restrict__ out) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < m2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(k / nmax), i = static_cast<int>(k % nmax);
    double v;
    if (i < nb[b] && j < nb[b])
      v = A[(static_cast<size_t>(f0[b]) + i) * n + f0[b] + j];
    else
      v = (i == j) ? pad : 0.0;
    out[b * m2 + k] = v;
  }
}

__global__ void unpack_kernel(const double* __restrict__ in, int nmax, const int* __restrict__ f0,
                              const int* __restrict__ nb, size_t n, double* __restrict__ A) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  const int nn = nb[b];
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       k < static_cast<size_t>(nn) * nn; k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(k / nn), i = static_cast<int>(k % nn);
    A[(static_cast<size_t>(f0[b]) + i) * n + f0[b] + j] = in[b * m2 + static_cast<size_t>(j) * nmax + i];
  }
}

__global__ void mask_cols_kernel(double* __restrict__ C, int nmax, const int* __restrict__ nocc) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < m2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    if (static_cast<int>(k / nmax) >= nocc[b]) C[b * m2 + k] = 0.0;
}

__global__ void inv_sqrt_cols_kernel(const double* __restrict__ U, const double* __restrict__ w,
                                     int nmax, double* __restrict__ out) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < m2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    out[b * m2 + k] = U[b * m2 + k] / sqrt(w[static_cast<size_t>(b) * nmax + k / nmax]);
}

__global__ __launch_bounds__(256) void energy_kernel(const double* __restrict__ D,
                                                     const double* __restrict__ H,
                                                     const double* __restrict__ F, size_t n,
                                                     const int* __restrict__ f0,
                                                     const int* __restrict__ nb,
                                                     double* __restrict__ out) {
  __shared__ double part[256];
  const int b = blockIdx.x;
  const int nn = nb[b];
  double s = 0.0;
  for (int k = threadIdx.x; k < nn * nn; k += blockDim.x) {
    const size_t idx = (static_cast<size_t>(f0[b]) + k / nn) * n + f0[b] + k % nn;
    s += D[idx] * (H[idx] + F[idx]);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = part[0];
}

__global__ void antisym_kernel(const double* __restrict__ A, double* __restrict__ E, int nmax) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < m2;
       k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t j = k / nmax, i = k % nmax;
    E[b * m2 + k] = A[b * m2 + k] - A[b * m2 + i * nmax + j];
  }
}

// pairs[p] = (offset of e, offset of f) in doubles; one workgroup per pair
__global__ __launch_bounds__(256) void dots_kernel(const double* __restrict__ base,
                                                   const long long* __restrict__ pairs,
                                                   size_t len, double* __restrict__ out) {
  __shared__ double part[256];
  const double* e = base + pairs[2 * blockIdx.x];
  const double* f = base + pairs[2 * blockIdx.x + 1];
  double s = 0.0;
  for (size_t k = threadIdx.x; k < len; k += blockDim.x) s += e[k] * f[k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(256) void maxabs_kernel(const double* __restrict__ E, int nmax,
                                                     double* __restrict__ out) {
  __shared__ double part[256];
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  const double* e = E + blockIdx.x * m2;
  double s = 0.0;
  for (size_t k = threadIdx.x; k < m2; k += blockDim.x) s = fmax(s, fabs(e[k]));
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

// out_b = sum_{i < m} c[b * k + i] Fs[i][b]; Fs[i] is a batch of padded blocks
__global__ void combine_kernel(const double* __restrict__ Fs, size_t batch_stride, int m, int k,
                               const double* __restrict__ c, int nmax, double* __restrict__ out) {
  const int b = blockIdx.y;
  const size_t m2 = static_cast<size_t>(nmax) * nmax;
  for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < m2;
       q += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double v = 0.0;
    for (int i = 0; i < m; ++i) v += c[static_cast<size_t>(b) * k + i] * Fs[i * batch_stride + b * m2 + q];
    out[b * m2 + q] = v;
  }
}

unsigned blocks_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 1024) g = 1024;
  return static_cast<unsigned>(g ? g : 1);
}

// D <- (1 - a_b) D + a_b D_old on molecule b's diagonal block (a_b = 0: kept);
// D_old is the padded copy (nmax x nmax per molecule, as batch_pack writes it).
__global__ void blend_kernel(double* __restrict__ D, const double* __restrict__ Dold, size_t n,
                             const int* __restrict__ f0, const int* __restrict__ nb,
                             const double* __restrict__ alpha, int nmax) {
  const int b = blockIdx.y;
  const double a = alpha[b];
  if (a == 0.0) return;
  const int nn = nb[b];
  const double* old = Dold + static_cast<size_t>(b) * nmax * nmax;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       k < static_cast<size_t>(nn) * nn; k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t i = k / nn, j = k % nn;
    const size_t ij = (f0[b] + i) * n + f0[b] + j;
    D[ij] = (1.0 - a) * D[ij] + a * old[j * nmax + i];
  }
}

// out + r * out_stride + s * nout^2 <- top-left nout x nout of
// in + r * in_stride + idx[s] * nin^2 (column-major blocks), r = blockIdx.z.
__global__ void gather_kernel(const double* __restrict__ in, size_t in_stride, int nin,
                              const int* __restrict__ idx, double* __restrict__ out,
                              size_t out_stride, int nout) {
  const int s = blockIdx.y, r = blockIdx.z;
  const double* src = in + r * in_stride + static_cast<size_t>(idx[s]) * nin * nin;
  double* dst = out + r * out_stride + static_cast<size_t>(s) * nout * nout;
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       k < static_cast<size_t>(nout) * nout; k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t j = k / nout, i = k % nout;
    dst[k] = src[j * nin + i];
  }
}

__global__ void fock_blocks_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                   const double* __restrict__ K, double* __restrict__ F, size_t n,
                                   const int* __restrict__ f0, const int* __restrict__ nb) {
  const int b = blockIdx.y, nn = nb[b];
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       k < static_cast<size_t>(nn) * nn; k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t ij = (f0[b] + k / nn) * n + f0[b] + k % nn;
    F[ij] = H[ij] + 2.0 * J[ij] - K[ij];
  }
}

}  // namespace

hipError_t batch_blend(double* D, const double* Dold, size_t n, const int* f0, const int* nb,
                       const double* alpha, int nmax, int nbatch, hipStream_t s) {
  if (nbatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(blend_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, D, Dold, n, f0, nb, alpha, nmax);
  return hipGetLastError();
}

hipError_t batch_gather(const double* in, size_t in_stride, int nin, const int* idx, int nout_batch,
                        int nring, double* out, size_t out_stride, int nout, hipStream_t s) {
  if (nout_batch <= 0 || nring <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel,
                     dim3(blocks_for(static_cast<size_t>(nout) * nout), nout_batch, nring),
                     dim3(256), 0, s, in, in_stride, nin, idx, out, out_stride, nout);
  return hipGetLastError();
}

hipError_t batch_fock(const double* H, const double* J, const double* K, double* F, size_t n,
                      const int* f0, const int* nb, int nmol, int nmax, hipStream_t s) {
  if (nmol <= 0) return hipSuccess;
  hipLaunchKernelGGL(fock_blocks_kernel,
                     dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nmol), dim3(256), 0, s,
                     H, J, K, F, n, f0, nb);
  return hipGetLastError();
}

hipError_t batch_pack(const double* A, size_t n, const int* f0, const int* nb, int nmax, int nbatch,
                      double pad, double* out, hipStream_t s) {
  hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, A, n, f0, nb, nmax, pad, out);
  return hipGetLastError();
}

hipError_t batch_unpack(const double* in, int nmax, const int* f0, const int* nb, int nbatch,
                        size_t n, double* A, hipStream_t s) {
  hipLaunchKernelGGL(unpack_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, in, nmax, f0, nb, n, A);
  return hipGetLastError();
}

hipError_t batch_mask_cols(double* C, int nmax, const int* nocc, int nbatch, hipStream_t s) {
  hipLaunchKernelGGL(mask_cols_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, C, nmax, nocc);
  return hipGetLastError();
}

hipError_t batch_inv_sqrt_cols(const double* U, const double* w, int nmax, int nbatch, double* out,
                               hipStream_t s) {
  hipLaunchKernelGGL(inv_sqrt_cols_kernel,
                     dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch), dim3(256), 0, s,
                     U, w, nmax, out);
  return hipGetLastError();
}

hipError_t batch_energy(const double* D, const double* H, const double* F, size_t n, const int* f0,
                        const int* nb, int nbatch, double* out, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel, dim3(nbatch), dim3(256), 0, s, D, H, F, n, f0, nb, out);
  return hipGetLastError();
}

hipError_t batch_antisym(const double* A, double* E, int nmax, int nbatch, hipStream_t s) {
  hipLaunchKernelGGL(antisym_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, A, E, nmax);
  return hipGetLastError();
}

hipError_t batch_dots(const double* base, const long long* pairs, int npairs, size_t len,
                      double* out, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(dots_kernel, dim3(npairs), dim3(256), 0, s, base, pairs, len, out);
  return hipGetLastError();
}

hipError_t batch_maxabs(const double* E, int nmax, int nbatch, double* out, hipStream_t s) {
  hipLaunchKernelGGL(maxabs_kernel, dim3(nbatch), dim3(256), 0, s, E, nmax, out);
  return hipGetLastError();
}

hipError_t batch_combine(const double* Fs, size_t batch_stride, int m, int k, const double* c,
                         int nmax, int nbatch, double* out, hipStream_t s) {
  hipLaunchKernelGGL(combine_kernel, dim3(blocks_for(static_cast<size_t>(nmax) * nmax), nbatch),
                     dim3(256), 0, s, Fs, batch_stride, m, k, c, nmax, out);
  return hipGetLastError();
}

}  // namespace rchem

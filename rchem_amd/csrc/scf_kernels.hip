// Small elementwise and reduction kernels of the SCF step (gfx950).
//   fill:          graph-safe clearing of accumulators (no hipMemsetAsync);
//   fock:          F = H + 2J - K (rchem.rs:110-127);
//   energy:        E = sum D o (H + F) (rchem.rs:106-108), one workgroup,
//                  fixed reduction order, so it is deterministic;
//   scale_columns: U diag(lambda^-1/2), the middle of Loewdin X (rchem.rs:44-51);
//   antisym:       E = A - A^T, the DIIS commutator FDS - SDF from A = FDS.
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

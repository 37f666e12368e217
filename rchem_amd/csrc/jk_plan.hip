// Device-side quartet planning for one Fock build (gfx950).
//
// The work of a build is a class-sorted queue of tiles (one bra shell pair x
// up to 64 ket pairs of one bucket).  It is rebuilt on the device for every
// density, with no host round trip, in four steps:
//   1. dmax:   max |D_ij| (order-preserving integer atomics on |D|'s bits);
//   2. plan:   one thread per task (bra pair, ket bucket): the ket range that
//              survives Schwarz screening Q_ab Q_cd >= tau (Q sorted descending
//              inside a bucket, so the survivors are a prefix found by binary
//              search), split into an FP64 prefix (Q_ab Q_cd max|D| >=
//              mixed_thresh) and an FP32 remainder; tile counts per task;
//   3. scan:   exclusive prefix sums of the tile counts (hipCUB);
//   4. expand: each task writes its tiles; per class, the [begin, end) tile
//              range of each precision is written for the persistent ERI
//              kernels, which read it from device memory.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "types.hpp"

namespace rchem {

__global__ void dmax_kernel(const double* __restrict__ D, size_t n2,
                            unsigned long long* __restrict__ out) {
  unsigned long long m = 0;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n2;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double v = fabs(D[i]);
    const unsigned long long b = __double_as_longlong(v);  // non-negative: bits order = value order
    m = b > m ? b : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// First index in [lo, hi) whose Q falls below `need` (Q non-increasing).
__device__ __forceinline__ int first_below(const double* __restrict__ Q, int lo, int hi,
                                           double need) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (Q[mid] >= need) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void plan_kernel(const PlanArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntasks) {
    if (t == a.ntasks) {  // the extra slot makes the exclusive scan's last entry the total
      a.cnt64[t] = 0;
      a.cnt32[t] = 0;
    }
    return;
  }
  // class of this task (few classes: linear search)
  int c = 0;
  while (c + 1 < a.nclasses && a.classes[c + 1].task_begin <= t) ++c;
  const PlanClass pc = a.classes[c];
  const int local = t - pc.task_begin;
  const int i = pc.bra_begin + local / pc.nbk;
  const int4 bk = a.buckets[pc.bk_off + local % pc.nbk];
  int lo = bk.x, hi = bk.y;
  if (pc.same) hi = (lo > i) ? lo : min(hi, i + 1);
  const double qi = a.Q[i];
  if (a.tau > 0.0) hi = (qi > 0.0) ? first_below(a.Q, lo, hi, a.tau / qi) : lo;
  int mid = hi;
  if (a.mixed_thresh > 0.0 && hi > lo) {
    const double dm = *a.dmax;
    const double need = (qi > 0.0 && dm > 0.0) ? a.mixed_thresh / (qi * dm) : HUGE_VAL;
    mid = first_below(a.Q, lo, hi, need);
  }
  a.lo[t] = lo;
  a.mid[t] = mid;
  a.hi[t] = hi;
  a.cnt64[t] = (mid - lo + 63) >> 6;
  a.cnt32[t] = (hi - mid + 63) >> 6;
  if (hi > lo) {
    const double npb = a.pairs[i].np, npk = bk.z;
    double* st = a.stats + 4 * c;
    if (mid > lo) {
      unsafeAtomicAdd(st + 0, double(mid - lo));
      unsafeAtomicAdd(st + 2, double(mid - lo) * npb * npk);
    }
    if (hi > mid) {
      unsafeAtomicAdd(st + 1, double(hi - mid));
      unsafeAtomicAdd(st + 3, double(hi - mid) * npb * npk);
    }
  }
}

// off64/off32: exclusive scans of the tile counts (length ntasks + 1).
__global__ void expand_kernel(const PlanArgs a, const int* __restrict__ off64,
                              const int* __restrict__ off32, Tile* __restrict__ tiles64,
                              Tile* __restrict__ tiles32, int cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntasks) return;
  int c = 0;
  while (c + 1 < a.nclasses && a.classes[c + 1].task_begin <= t) ++c;
  const PlanClass pc = a.classes[c];
  const int i = pc.bra_begin + (t - pc.task_begin) / pc.nbk;
  const int lo = a.lo[t], mid = a.mid[t], hi = a.hi[t];
  int o = off64[t];
  for (int k = lo; k < mid && o < cap; k += 64, ++o) tiles64[o] = Tile{i, k, min(64, mid - k), 0};
  o = off32[t];
  for (int k = mid; k < hi && o < cap; k += 64, ++o) tiles32[o] = Tile{i, k, min(64, hi - k), 0};
}

// ranges[4c .. 4c+3] = {begin64, end64, begin32, end32} of class c; the end of
// the last class is the scan total, clamped to the buffer capacity.
__global__ void ranges_kernel(const PlanArgs a, const int* __restrict__ off64,
                              const int* __restrict__ off32, int* __restrict__ ranges, int cap) {
  const int c = threadIdx.x;
  if (c >= a.nclasses) return;
  const int tb = a.classes[c].task_begin;
  const int te = (c + 1 < a.nclasses) ? a.classes[c + 1].task_begin : a.ntasks;
  ranges[4 * c + 0] = min(off64[tb], cap);
  ranges[4 * c + 1] = min(off64[te], cap);
  ranges[4 * c + 2] = min(off32[tb], cap);
  ranges[4 * c + 3] = min(off32[te], cap);
}

hipError_t dmax_launch(const double* D, size_t n2, unsigned long long* out, hipStream_t s) {
  size_t blocks = (n2 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(dmax_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, D, n2,
                     out);
  return hipGetLastError();
}

// Temporary-storage bytes of the scans over n + 1 entries.
size_t plan_scan_bytes(int ntasks) {
  size_t bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<int*>(nullptr),
                                       static_cast<int*>(nullptr), ntasks + 1) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t plan_launch(const PlanArgs& a, int* off64, int* off32, Tile* tiles64, Tile* tiles32,
                       int cap, int* ranges, void* scan_tmp, size_t scan_bytes, hipStream_t s) {
  const unsigned blocks = static_cast<unsigned>((a.ntasks + 1 + 255) / 256);
  hipLaunchKernelGGL(plan_kernel, dim3(blocks), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t b = scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, a.cnt64, off64, a.ntasks + 1, s);
  if (e != hipSuccess) return e;
  b = scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, a.cnt32, off32, a.ntasks + 1, s);
  if (e != hipSuccess) return e;
  if (a.ntasks > 0) {
    hipLaunchKernelGGL(expand_kernel, dim3(static_cast<unsigned>((a.ntasks + 255) / 256)),
                       dim3(256), 0, s, a, off64, off32, tiles64, tiles32, cap);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(ranges_kernel, dim3(1), dim3(64), 0, s, a, off64, off32, ranges, cap);
  return hipGetLastError();
}

}  // namespace rchem

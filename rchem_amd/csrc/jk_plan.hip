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
//   4. expand: each task writes its tiles;
//   5. chunks: the tiles of one (class, bra pair) unit are cut into chunks of
//              at most the class's chunk size per precision (from the tuning
//              table; a workgroup's unit of work when K is
//              digested in LDS), counted, scanned and written (optionally
//              radix-sorted by first ket pair within each class); per class the
//              [begin, end) tile and chunk ranges of each precision go to device
//              memory, where the persistent ERI kernels read them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "types.hpp"

namespace rchem {

// max over the block (256 threads), the result in thread 0
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long m) {
  __shared__ unsigned long long part[4];
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < 4; ++w) m = part[w] > m ? part[w] : m;
  return m;
}

// One atomic per workgroup (of at most 256): all of them meet on one address.
__global__ __launch_bounds__(256) void dmax_kernel(const double* __restrict__ D, size_t n2,
                                                   unsigned long long* __restrict__ out) {
  unsigned long long m = 0;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n2;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double v = fabs(D[i]);
    const unsigned long long b = __double_as_longlong(v);  // non-negative: bits order = value order
    m = b > m ? b : m;
  }
  m = block_max_u64(m);
  if (threadIdx.x == 0) atomicMax(out, m);
}

// The same over the diagonal blocks [f0_m, f0_m + nb_m)^2 of a multi-molecule
// basis (the rest of D is never read).
__global__ void dmax_blocks_kernel(const double* __restrict__ D, size_t n,
                                   const int* __restrict__ f0, const int* __restrict__ nb,
                                   unsigned long long* __restrict__ out) {
  const int m = blockIdx.y, b = nb[m];
  const size_t o = f0[m];
  unsigned long long mx = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < b * b; k += gridDim.x * blockDim.x) {
    const double v = fabs(D[(o + k / b) * n + o + k % b]);
    const unsigned long long bits = __double_as_longlong(v);
    mx = bits > mx ? bits : mx;
  }
  mx = block_max_u64(mx);
  if (threadIdx.x == 0) atomicMax(out, mx);
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

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int class_of_task(const PlanArgs& a, int t) {
  int c = 0;
  while (c + 1 < a.nclasses && a.classes[c + 1].task_begin <= t) ++c;
  return c;
}

__global__ void plan_kernel(const PlanArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < a.ntasks;
  if (t == a.ntasks) {  // the extra slot makes the exclusive scan's last entry the total
    a.cnt64[t] = 0;
    a.cnt32[t] = 0;
  }
  int c = 0;
  double q64 = 0.0, q32 = 0.0, p64 = 0.0, p32 = 0.0;
  if (valid) {
    c = class_of_task(a, t);
    const PlanClass pc = a.classes[c];
    const int local = t - pc.task_begin;
    const int i = pc.bra_begin + local / pc.nbk;
    const int4 bk = a.buckets[pc.bk_off + local % pc.nbk];
    int lo = bk.x, hi = bk.y;
    bool all64 = false;
    if (a.segs) {  // the bra's molecule only
      const int mol = a.pair_mol[i];
      const int2 sg = a.segs[static_cast<size_t>(pc.bk_off + local % pc.nbk) * a.nmol + mol];
      const int state = a.mol_active ? a.mol_active[mol] : 1;
      lo = sg.x;
      hi = state == 0 ? sg.x : sg.y;  // a retired molecule: none
      all64 = state == 2;             // a molecule past its precision switch: no FP32
    }
    if (pc.same) hi = (lo > i) ? lo : min(hi, i + 1);
    const double qi = a.Q[i];
    if (a.tau > 0.0) hi = (qi > 0.0) ? first_below(a.Q, lo, hi, a.tau / qi) : lo;
    int mid = hi;
    if (a.mixed_thresh > 0.0 && hi > lo && !all64) {
      const double dm = *a.dmax;
      const double need = (qi > 0.0 && dm > 0.0) ? a.mixed_thresh / (qi * dm) : HUGE_VAL;
      mid = first_below(a.Q, lo, hi, need);
    }
    a.lo[t] = lo;
    a.mid[t] = mid;
    a.hi[t] = hi;
    a.cnt64[t] = (mid - lo + 63) >> 6;
    a.cnt32[t] = (hi - mid + 63) >> 6;
    const double npp = static_cast<double>(a.pairs[i].np) * bk.z;
    q64 = mid - lo;
    q32 = hi - mid;
    p64 = q64 * npp;
    p32 = q32 * npp;
  }
  // statistics: one atomic per wave when the wave's tasks share a class, into
  // the wave's replica
  double* const stats = a.stats + static_cast<size_t>((t >> 6) & (kPlanStatsReplicas - 1)) *
                                      a.stats_stride;
  const int c0 = __shfl(c, 0, 64);
  const int lastv = __shfl(valid ? 1 : 0, 63, 64);
  const int c63 = __shfl(c, 63, 64);
  if (lastv && c0 == c63) {
    q64 = wave_sum_d(q64);
    q32 = wave_sum_d(q32);
    p64 = wave_sum_d(p64);
    p32 = wave_sum_d(p32);
    if ((threadIdx.x & 63) == 0) {
      double* st = stats + 4 * c0;
      if (q64 > 0) { unsafeAtomicAdd(st + 0, q64); unsafeAtomicAdd(st + 2, p64); }
      if (q32 > 0) { unsafeAtomicAdd(st + 1, q32); unsafeAtomicAdd(st + 3, p32); }
    }
  } else if (valid) {
    double* st = stats + 4 * c;
    if (q64 > 0) { unsafeAtomicAdd(st + 0, q64); unsafeAtomicAdd(st + 2, p64); }
    if (q32 > 0) { unsafeAtomicAdd(st + 1, q32); unsafeAtomicAdd(st + 3, p32); }
  }
}

// off64/off32: exclusive scans of the tile counts (length ntasks + 1).
__global__ void expand_kernel(const PlanArgs a, const int* __restrict__ off64,
                              const int* __restrict__ off32, Tile* __restrict__ tiles64,
                              Tile* __restrict__ tiles32, int cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntasks) return;
  const PlanClass pc = a.classes[class_of_task(a, t)];
  const int i = pc.bra_begin + (t - pc.task_begin) / pc.nbk;
  const int lo = a.lo[t], mid = a.mid[t], hi = a.hi[t];
  int o = off64[t];
  for (int k = lo; k < mid && o < cap; k += 64, ++o) tiles64[o] = Tile{i, k, min(64, mid - k), 0};
  o = off32[t];
  for (int k = mid; k < hi && o < cap; k += 64, ++o) tiles32[o] = Tile{i, k, min(64, hi - k), 0};
}

__device__ __forceinline__ int class_of_unit(const PlanArgs& a, int u) {
  int c = 0;
  while (c + 1 < a.nclasses && a.classes[c + 1].unit_begin <= u) ++c;
  return c;
}

// Unit u = (class c, bra pair): its tasks are [t0, t0 + nbk).
__device__ __forceinline__ void unit_tasks(const PlanArgs& a, int u, int* c, int* bra, int* t0,
                                           int* t1) {
  *c = class_of_unit(a, u);
  const PlanClass pc = a.classes[*c];
  const int local = u - pc.unit_begin;
  *bra = pc.bra_begin + local;
  *t0 = pc.task_begin + local * pc.nbk;
  *t1 = *t0 + pc.nbk;
}

__global__ void chunk_count_kernel(const PlanArgs a, const int* __restrict__ off64,
                                   const int* __restrict__ off32, int* __restrict__ chc64,
                                   int* __restrict__ chc32) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > a.nunits) return;
  if (u == a.nunits) {
    chc64[u] = 0;
    chc32[u] = 0;
    return;
  }
  int c, bra, t0, t1;
  unit_tasks(a, u, &c, &bra, &t0, &t1);
  const int ch64 = a.classes[c].chunk64, ch32 = a.classes[c].chunk32;
  chc64[u] = (off64[t1] - off64[t0] + ch64 - 1) / ch64;
  chc32[u] = (off32[t1] - off32[t0] + ch32 - 1) / ch32;
}

// cap bounds the tile offsets (tile buffers), ccap the chunk records written
// (chunk buffers); every record below min(chunk count, ccap) is written.
// keys64 / keys32 (when set): per chunk, (c << key_shift) | (first ket pair of
// the chunk / 64), c the chunk's class (segment) in this pass.
__global__ void chunk_expand_kernel(const PlanArgs a, const int* __restrict__ off64,
                                    const int* __restrict__ off32, const int* __restrict__ choff64,
                                    const int* __restrict__ choff32, Chunk* __restrict__ chunks64,
                                    Chunk* __restrict__ chunks32, int cap, int ccap,
                                    const Tile* __restrict__ tiles64,
                                    const Tile* __restrict__ tiles32, unsigned* __restrict__ keys64,
                                    unsigned* __restrict__ keys32, int key_shift) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= a.nunits) return;
  int c, bra, t0, t1;
  unit_tasks(a, u, &c, &bra, &t0, &t1);
  const unsigned khigh = static_cast<unsigned>(c) << key_shift;
  // a class planned in bra order keeps it: its keys differ only by class, and
  // the sort is stable
  const int order = a.classes[c].order;
  {
    const int ch = a.classes[c].chunk64;
    const int tb = min(off64[t0], cap), te = min(off64[t1], cap);
    int o = choff64[u];
    for (int k = tb; k < te && o < ccap; k += ch, ++o) {
      chunks64[o] = Chunk{bra, k, min(k + ch, te), 0};
      if (keys64) keys64[o] = khigh | ((order & 1) ? static_cast<unsigned>(tiles64[k].ket0 >> 6) : 0u);
    }
  }
  {
    const int ch = a.classes[c].chunk32;
    const int tb = min(off32[t0], cap), te = min(off32[t1], cap);
    int o = choff32[u];
    for (int k = tb; k < te && o < ccap; k += ch, ++o) {
      chunks32[o] = Chunk{bra, k, min(k + ch, te), 0};
      if (keys32) keys32[o] = khigh | ((order & 2) ? static_cast<unsigned>(tiles32[k].ket0 >> 6) : 0u);
    }
  }
}

// keys[i] = all ones for i in [total, n): the entries past the chunk count
// sort to the end.
__global__ void pad_keys_kernel(unsigned* __restrict__ keys, const int* __restrict__ total, int n) {
  const int t = min(*total, n);
  for (int i = t + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    keys[i] = ~0u;
}

// ranges[8c .. 8c+7] = tiles {b64, e64, b32, e32}, chunks {b64, e64, b32, e32}
// of class c, clamped to the buffer capacity.
__global__ void ranges_kernel(const PlanArgs a, const int* __restrict__ off64,
                              const int* __restrict__ off32, const int* __restrict__ choff64,
                              const int* __restrict__ choff32, int* __restrict__ ranges, int cap,
                              int ccap) {
  const int c = threadIdx.x;
  if (c >= a.nclasses) return;
  const int tb = a.classes[c].task_begin;
  const int te = (c + 1 < a.nclasses) ? a.classes[c + 1].task_begin : a.ntasks;
  const int ub = a.classes[c].unit_begin;
  const int ue = (c + 1 < a.nclasses) ? a.classes[c + 1].unit_begin : a.nunits;
  ranges[8 * c + 0] = min(off64[tb], cap);
  ranges[8 * c + 1] = min(off64[te], cap);
  ranges[8 * c + 2] = min(off32[tb], cap);
  ranges[8 * c + 3] = min(off32[te], cap);
  ranges[8 * c + 4] = min(choff64[ub], ccap);
  ranges[8 * c + 5] = min(choff64[ue], ccap);
  ranges[8 * c + 6] = min(choff32[ub], ccap);
  ranges[8 * c + 7] = min(choff32[ue], ccap);
}

hipError_t dmax_launch(const double* D, size_t n2, unsigned long long* out, hipStream_t s) {
  size_t blocks = (n2 + 255) / 256;
  if (blocks > 512) blocks = 512;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(dmax_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, D, n2,
                     out);
  return hipGetLastError();
}

// Bits of the class field of a chunk key (classes of one pass).
int plan_class_bits(int nclasses) {
  int bits = 0;
  while ((1 << bits) < nclasses) ++bits;
  return bits;
}

// Temporary-storage bytes of a radix sort of n chunk keys and chunks.
size_t plan_sort_bytes(int n) {
  size_t bytes = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<unsigned*>(nullptr),
                                         static_cast<unsigned*>(nullptr),
                                         static_cast<Chunk*>(nullptr),
                                         static_cast<Chunk*>(nullptr), n, 0, 32) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t dmax_blocks_launch(const double* D, size_t n, const int* f0, const int* nb, int nmol,
                              int nmax, unsigned long long* out, hipStream_t s) {
  if (nmol <= 0) return hipSuccess;
  const unsigned gx = static_cast<unsigned>(min(64, (nmax * nmax + 255) / 256));
  hipLaunchKernelGGL(dmax_blocks_kernel, dim3(gx, nmol), dim3(256), 0, s, D, n, f0, nb, out);
  return hipGetLastError();
}

// Temporary-storage bytes of a scan over n + 1 entries.
size_t plan_scan_bytes(int n) {
  size_t bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<int*>(nullptr),
                                       static_cast<int*>(nullptr), n + 1) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t plan_launch(const PlanArgs& a, const PlanBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(plan_kernel, dim3(static_cast<unsigned>((a.ntasks + 1 + 255) / 256)),
                     dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t sb = b.scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, sb, a.cnt64, b.off64, a.ntasks + 1, s);
  if (e != hipSuccess) return e;
  sb = b.scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, sb, a.cnt32, b.off32, a.ntasks + 1, s);
  if (e != hipSuccess) return e;
  if (a.ntasks > 0) {
    hipLaunchKernelGGL(expand_kernel, dim3(static_cast<unsigned>((a.ntasks + 255) / 256)),
                       dim3(256), 0, s, a, b.off64, b.off32, b.tiles64, b.tiles32, b.cap);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(chunk_count_kernel, dim3(static_cast<unsigned>((a.nunits + 1 + 255) / 256)),
                     dim3(256), 0, s, a, b.off64, b.off32, b.chc64, b.chc32);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  sb = b.scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, sb, b.chc64, b.choff64, a.nunits + 1, s);
  if (e != hipSuccess) return e;
  sb = b.scan_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, sb, b.chc32, b.choff32, a.nunits + 1, s);
  if (e != hipSuccess) return e;
  const bool sorting = b.key_bits > 0 && b.ccap > 0;
  const bool sort64 = sorting && (b.sort_mask & 1);
  // FP64-only builds have no FP32 chunks
  const bool sort32 = sorting && (b.sort_mask & 2) && a.mixed_thresh > 0.0;
  // chunk records per precision: every record below min(count, ccap) is written
  const int ccap = sorting ? min(b.cap, b.ccap) : b.cap;
  Chunk* out64 = sort64 ? b.chunks_raw : b.chunks64;
  Chunk* out32 = sort32 ? b.chunks_raw + b.ccap : b.chunks32;
  unsigned* k64 = sort64 ? b.keys_raw : nullptr;
  unsigned* k32 = sort32 ? b.keys_raw + b.ccap : nullptr;
  if (a.nunits > 0) {
    hipLaunchKernelGGL(chunk_expand_kernel, dim3(static_cast<unsigned>((a.nunits + 255) / 256)),
                       dim3(256), 0, s, a, b.off64, b.off32, b.choff64, b.choff32, out64, out32,
                       b.cap, ccap, b.tiles64, b.tiles32, k64, k32,
                       b.key_bits - plan_class_bits(a.nclasses));
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int p = 0; p < 2; ++p) {
    if (!(p ? sort32 : sort64)) continue;
    unsigned* kin = b.keys_raw + static_cast<size_t>(p) * b.ccap;
    const unsigned pad_blocks = static_cast<unsigned>(min((b.ccap + 255) / 256, 1024));
    hipLaunchKernelGGL(pad_keys_kernel, dim3(pad_blocks), dim3(256), 0, s, kin,
                       (p ? b.choff32 : b.choff64) + a.nunits, b.ccap);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t sb = b.sort_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(b.sort_tmp, sb, kin, b.keys_sorted,
                                           b.chunks_raw + static_cast<size_t>(p) * b.ccap,
                                           p ? b.chunks32 : b.chunks64, b.ccap, 0, b.key_bits, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(ranges_kernel, dim3(1), dim3(64), 0, s, a, b.off64, b.off32, b.choff64,
                     b.choff32, b.ranges, b.cap, ccap);
  return hipGetLastError();
}

}  // namespace rchem

// Wave-cooperative contracted ERIs for the classes whose integrals do not fit
// one lane's registers (total angular momentum L >= 5: (dp|pp) ... (dd|dd)),
// and for any class on request.
//
// One wavefront (a 64-thread workgroup) evaluates one contracted shell
// quartet at a time:
//   1. lanes set up a batch of up to kGenericPqBatch primitive quartets in
//      parallel (exponent factors, PA, WP, QC, WQ and G_m = K_ab K_cd
//      (zeta+eta)^(-1/2) F_m(T), m = 0..L), one record each, into LDS;
//   2. per primitive quartet the Head-Gordon--Pople vertical recurrence runs
//      as the class's operation table: operations of one dependency level are
//      independent and spread over the 64 lanes, a level boundary is a
//      workgroup barrier (one wave), and within a level the operations come in
//      groups of one (side, axis), whose recurrence coefficients are
//      wave-uniform.  Target integrals [e0|f0]^(0) add into the contracted
//      accumulators as they are produced;
//   3. after the primitive loop the horizontal recurrence runs the same way on
//      the accumulators, which are the first slots of the same buffer;
//   4. lanes take the class's integrals in canonical component order for the
//      output mode (shell-quartet blocks, dense tensor, Schwarz diagonal, or
//      J/K digestion with FP64 atomics and the 8-fold weights).
// The recurrences are the straight-line generator's (same nodes, same axis
// choices), so both engines agree to rounding; the reference's HGP is
// chgp.c:44-135 and chgp.c:412-586.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "boys.hpp"
#include "types.hpp"

namespace rchem {

namespace {

// Record fields of one primitive quartet
constexpr int kPA = 0, kWP = 3, kQC = 6, kWQ = 9, kOo2z = 12, kOo2e = 13, kOo2ze = 14,
              kRoz = 15, kRoe = 16, kG = 17;

__device__ __forceinline__ int rec_size(int L) { return kG + L + 1; }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int BM>
__device__ void setup_record(const EriArgs& a, const PairRec& bp, const PairRec& kp, int ib,
                             int ik, int L, double* __restrict__ rec) {
  const PrimRec pb = a.prims[bp.p0 + ib * bp.ps];
  const PrimRec pk = a.prims[kp.p0 + ik * kp.ps];
  const double zb = pb.zeta, zk = pk.zeta;
  const double rs = fast_rsqrt(zb + zk);
  const double ooze = rs * rs;
  const double roz = zk * ooze, roe = zb * ooze;
  const double PQx = pb.px - pk.px, PQy = pb.py - pk.py, PQz = pb.pz - pk.pz;
  const double T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
  rec[kPA + 0] = pb.px - bp.A[0];
  rec[kPA + 1] = pb.py - bp.A[1];
  rec[kPA + 2] = pb.pz - bp.A[2];
  rec[kWP + 0] = -roz * PQx;
  rec[kWP + 1] = -roz * PQy;
  rec[kWP + 2] = -roz * PQz;
  rec[kQC + 0] = pk.px - kp.A[0];
  rec[kQC + 1] = pk.py - kp.A[1];
  rec[kQC + 2] = pk.pz - kp.A[2];
  rec[kWQ + 0] = roe * PQx;
  rec[kWQ + 1] = roe * PQy;
  rec[kWQ + 2] = roe * PQz;
  rec[kOo2z] = pb.oo2z;
  rec[kOo2e] = pk.oo2z;
  rec[kOo2ze] = 0.5 * ooze;
  rec[kRoz] = roz;
  rec[kRoe] = roe;
  double* G = rec + kG;
  if constexpr (BM == kBoysCompat) {
    for (int m = 0; m <= L; ++m) G[m] = fgamma_compat(static_cast<double>(m), T);
  } else {
    boys_exact_rt(L, T, a.boys_table, G);
  }
  const double pref = pb.kab * pk.kab * rs;
  for (int m = 0; m <= L; ++m) G[m] *= pref;
}

// One contracted quartet (bra pair bidx, ket pair kidx) into H[out_src[k]].
template <int BM>
__device__ void quartet(const EriArgs& a, const GenericProgram& p, int bidx, int kidx,
                        double* __restrict__ pq, double* __restrict__ R, int lane) {
  const PairRec bp = a.pairs[bidx];
  const PairRec kp = a.pairs[kidx];
  const int L = p.L, RS = rec_size(L);
  __syncthreads();  // the previous quartet's integrals have been read
  double* __restrict__ acc = R;
  double* __restrict__ V = R + p.nacc;
  for (int k = lane; k < p.nacc; k += 64) acc[k] = 0.0;
  if (lane == 0) V[0] = 0.0;
  const int nk = kp.np, npq = bp.np * kp.np;
  for (int b0 = 0; b0 < npq; b0 += kGenericPqBatch) {
    const int nb = min(kGenericPqBatch, npq - b0);
    __syncthreads();  // the previous batch's records are consumed
    if (lane < nb) {
      const int j = b0 + lane;
      setup_record<BM>(a, bp, kp, j / nk, j % nk, L, pq + lane * RS);
    }
    __syncthreads();
    for (int q = 0; q < nb; ++q) {
      const double* __restrict__ rec = pq + q * RS;
      if (lane <= L) V[1 + lane] = rec[kG + lane];
      __syncthreads();
      for (int g = 0; g < p.nvgroups; ++g) {
        const uint4 gr = p.vgroups[g];
        const int side = gr.z & 1, ax = (gr.z >> 1) & 3;
        const double c0 = rec[(side ? kQC : kPA) + ax];
        const double c1 = rec[(side ? kWQ : kWP) + ax];
        const double o = rec[side ? kOo2e : kOo2z];
        const double r = rec[side ? kRoe : kRoz];
        const double oz = rec[kOo2ze];
        for (int op = static_cast<int>(gr.x) + lane; op < static_cast<int>(gr.y); op += 64) {
          const uint4 w = p.vops[op];
          const double n1 = static_cast<double>((w.w >> 16) & 15u);
          const double n2 = static_cast<double>((w.w >> 20) & 15u);
          double v = c0 * V[w.x >> 16] + c1 * V[w.y & 0xffffu];
          v += (n1 * o) * (V[w.y >> 16] - r * V[w.z & 0xffffu]);
          v += (n2 * oz) * V[w.z >> 16];
          V[w.x & 0xffffu] = v;
          const unsigned ai = w.w & 0xffffu;
          if (ai != 0xffffu) acc[ai] += v;
        }
        if (gr.z & 8u) __syncthreads();
      }
      if (p.nvgroups == 0 && lane == 0) acc[0] += V[1];  // (ss|ss): the target is G_0
    }
  }
  __syncthreads();
  // horizontal recurrence on the contracted accumulators
  double* __restrict__ H = R;
  for (int g = 0; g < p.nhgroups; ++g) {
    const uint4 gr = p.hgroups[g];
    const int side = gr.z & 1, ax = (gr.z >> 1) & 3;
    const double xb = ax == 0 ? bp.AB[0] : (ax == 1 ? bp.AB[1] : bp.AB[2]);
    const double xk = ax == 0 ? kp.AB[0] : (ax == 1 ? kp.AB[1] : kp.AB[2]);
    const double X = side ? xk : xb;
    for (int op = static_cast<int>(gr.x) + lane; op < static_cast<int>(gr.y); op += 64) {
      const uint2 w = p.hops[op];
      H[w.x & 0xffffu] = H[w.x >> 16] + X * H[w.y];
    }
    if (gr.z & 8u) __syncthreads();
  }
  __syncthreads();
}

__device__ __forceinline__ int ncart_rt(int l) { return (l + 1) * (l + 2) / 2; }

template <int MODE>
__device__ void epilogue(const EriArgs& a, const GenericProgram& p, const Tile& t, int bidx,
                         int kidx, const double* __restrict__ H, int lane) {
  const PairRec bp = a.pairs[bidx];
  const PairRec kp = a.pairs[kidx];
  const int NA = ncart_rt(p.la), NB = ncart_rt(p.lb), NC = ncart_rt(p.lc), ND = ncart_rt(p.ld);
  const size_t n = static_cast<size_t>(a.nbf);
  const int fa = bp.fa, fb = bp.fb, fc = kp.fa, fd = kp.fb;
  if constexpr (MODE == kModeQuartet) {
    for (int k = lane; k < p.nout; k += 64)
      a.out[static_cast<size_t>(t.pad) + k] = p.out_norm[k] * H[p.out_src[k]];
  } else if constexpr (MODE == kModeDiag) {
    double m = 0.0;
    for (int k = lane; k < NA * NB; k += 64) {
      const int ia = k / NB, ib = k % NB;
      const int o = ((ia * NB + ib) * NC + ia) * ND + ib;
      m = fmax(m, fabs(p.out_norm[o] * H[p.out_src[o]]));
    }
    m = wave_max(m);
    if (lane == 0) a.out[kidx] = sqrt(m);
  } else if constexpr (MODE == kModeTensor) {
    for (int k = lane; k < p.nout; k += 64) {
      const double v = p.out_norm[k] * H[p.out_src[k]];
      const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
      const size_t P = fa + ia, Q = fb + ib, Rr = fc + ic, S = fd + id;
      double* I = a.out;
      I[((P * n + Q) * n + Rr) * n + S] = v;
      I[((Q * n + P) * n + Rr) * n + S] = v;
      I[((P * n + Q) * n + S) * n + Rr] = v;
      I[((Q * n + P) * n + S) * n + Rr] = v;
      I[((Rr * n + S) * n + P) * n + Q] = v;
      I[((S * n + Rr) * n + P) * n + Q] = v;
      I[((Rr * n + S) * n + Q) * n + P] = v;
      I[((S * n + Rr) * n + Q) * n + P] = v;
    }
  } else {
    // 8-fold weight: each unique (ab|cd) stands for deg ordered quartets; the
    // lower-half accumulators get deg/8 of every permutation's contribution
    // and J + J^T, K + K^T restore the rest
    double s = 0.125;
    if (bp.flags & 1) s *= 2.0;
    if (kp.flags & 1) s *= 2.0;
    if (bidx != kidx) s *= 2.0;
    const double* __restrict__ D = a.D;
    for (int k = lane; k < p.nout; k += 64) {
      const double v = s * p.out_norm[k] * H[p.out_src[k]];
      const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
      const size_t A = fa + ia, B = fb + ib, C = fc + ic, Dd = fd + id;
      unsafeAtomicAdd(&a.J[A * n + B], 2.0 * v * D[C * n + Dd]);
      unsafeAtomicAdd(&a.J[C * n + Dd], 2.0 * v * D[A * n + B]);
      unsafeAtomicAdd(&a.K[A * n + C], v * D[B * n + Dd]);
      unsafeAtomicAdd(&a.K[A * n + Dd], v * D[B * n + C]);
      unsafeAtomicAdd(&a.K[B * n + C], v * D[A * n + Dd]);
      unsafeAtomicAdd(&a.K[B * n + Dd], v * D[A * n + C]);
    }
  }
}

template <int MODE, int BM>
__global__ __launch_bounds__(64) void eri_generic_kernel(const EriArgs a, const GenericProgram p) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  double* pq = lds;
  double* R = lds + kGenericPqBatch * rec_size(p.L);
  int begin, end;
  if (a.tile_range) {
    begin = a.tile_range[0];
    end = a.tile_range[1];
  } else {
    begin = 0;
    end = a.ntiles;
  }
  // every workgroup leaves once its next tile index passes the end
  for (int ti = begin + static_cast<int>(blockIdx.x); ti < end; ti += gridDim.x) {
    const Tile t = a.tiles[ti];
    for (int i = 0; i < t.count; ++i) {
      const int kidx = t.ket0 + i;
      const int bidx = (MODE == kModeDiag) ? kidx : t.bra;
      quartet<BM>(a, p, bidx, kidx, pq, R, lane);
      epilogue<MODE>(a, p, t, bidx, kidx, R, lane);
    }
  }
}

template <int MODE, int BM>
hipError_t launch_mode(const EriArgs& a, const GenericProgram& p, unsigned grid, size_t lds,
                       hipStream_t s) {
  hipLaunchKernelGGL((eri_generic_kernel<MODE, BM>), dim3(grid), dim3(64), lds, s, a, p);
  return hipGetLastError();
}

}  // namespace

// LDS bytes of one workgroup for program p.
size_t eri_generic_lds_bytes(const GenericProgram& p) {
  const size_t rs = 17 + p.L + 1;
  const size_t buf = std::max(static_cast<size_t>(p.nacc + p.nv), static_cast<size_t>(p.nh));
  return sizeof(double) * (kGenericPqBatch * rs + buf);
}

// blocks: with a device tile range, the number of 4-wave blocks the per-lane
// kernels would use (this kernel launches 4 one-wave workgroups per block);
// without one, one workgroup per tile.
hipError_t eri_generic_launch(const EriArgs& a, const GenericProgram& p, int mode, int bm,
                              int blocks, hipStream_t s) {
  unsigned grid;
  if (a.tile_range) {
    if (blocks <= 0) return hipSuccess;
    grid = static_cast<unsigned>(blocks) * 4u;
  } else {
    if (a.ntiles <= 0) return hipSuccess;
    grid = static_cast<unsigned>(a.ntiles);
  }
  const size_t lds = eri_generic_lds_bytes(p);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  const bool compat = bm == kBoysCompat;
  switch (mode) {
    case kModeJK:
      return compat ? launch_mode<kModeJK, kBoysCompat>(a, p, grid, lds, s)
                    : launch_mode<kModeJK, kBoysExact>(a, p, grid, lds, s);
    case kModeDiag:
      return compat ? launch_mode<kModeDiag, kBoysCompat>(a, p, grid, lds, s)
                    : launch_mode<kModeDiag, kBoysExact>(a, p, grid, lds, s);
    case kModeTensor:
      return compat ? launch_mode<kModeTensor, kBoysCompat>(a, p, grid, lds, s)
                    : launch_mode<kModeTensor, kBoysExact>(a, p, grid, lds, s);
    case kModeQuartet:
      return compat ? launch_mode<kModeQuartet, kBoysCompat>(a, p, grid, lds, s)
                    : launch_mode<kModeQuartet, kBoysExact>(a, p, grid, lds, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace rchem

// Wave-cooperative contracted ERIs for the classes whose integrals do not fit
// one lane's registers (total angular momentum L >= 5: (dp|pp) ... (dd|dd)),
// for any class on request, and the derivative integrals of the nuclear
// gradient.
//
// One wavefront (a 64-thread workgroup) evaluates one contracted shell
// quartet at a time, P primitive quartets per batch (P a power of two, the
// class's table value or less for short contractions):
//   1. lanes 0..P-1 set up the batch's primitive-quartet records (exponent
//      factors, PA, WP, QC, WQ, 2 alpha of three centres and
//      G_m = K_ab K_cd (zeta+eta)^(-1/2) F_m(T), m = 0..L) field-major in LDS;
//   2. the Head-Gordon--Pople vertical recurrence runs as the class's
//      operation table on V[slot][q]: lane l always serves primitive quartet
//      q = l mod P and takes operations l / P, l / P + 64 / P, ... of a group,
//      a group being operations of one dependency level and one (side, axis)
//      whose recurrence coefficients the lane loads once; a level boundary is a
//      workgroup barrier (one wave);
//   3. the accumulation list adds each target [e0|f0]^(0), summed over the
//      batch with its set's weight (1 or 2 alpha of a centre), into the
//      contracted accumulators, one lane per accumulator;
//   4. after the primitive loop the horizontal recurrence runs the same way on
//      the accumulators (the first slots of the same buffer);
//   5. lanes take the outputs norm (H[plus] - n H[minus]) for the mode:
//      shell-quartet blocks, dense tensor, Schwarz diagonal, J/K digestion
//      with FP64 atomics and the 8-fold weights, or the gradient contraction.
// The recurrences are the straight-line generator's (same nodes, same axis
// choices); the reference's HGP is chgp.c:44-135 and chgp.c:412-586.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "boys.hpp"
#include "types.hpp"

namespace rchem {

namespace {

// Record fields of one primitive quartet (stored field-major: rec[f * P + q])
constexpr int kPA = 0, kWP = 3, kQC = 6, kWQ = 9, kOo2z = 12, kOo2e = 13, kOo2ze = 14,
              kRoz = 15, kRoe = 16, kWA = 17, kWB = 18, kWC = 19, kG = 20;

__device__ __forceinline__ int rec_fields(int L) { return kG + L + 1; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// pref * F_m(T), m = 0..M, into out[m * stride]: the exact Boys function of
// boys_exact_rt (Taylor row of F_M on the T0 grid, exp row, downward
// recursion; asymptote and upward recursion for T >= 36).
__device__ void boys_exact_strided(int M, double T, const double* __restrict__ table, double pref,
                                   double* __restrict__ out, int stride) {
  if (T < kBoysTMax) {
    const int k = static_cast<int>(T * kBoysInvH + 0.5);
    const double d = k * (1.0 / kBoysInvH) - T;
    const double* c = table + M * kBoysBlock + k * kBoysTerms;
    double f = c[kBoysTerms - 1];
    for (int j = kBoysTerms - 2; j >= 0; --j) f = c[j] + d * f;
    double pe = 1.0;
    for (int j = kBoysTerms - 1; j >= 1; --j) pe = 1.0 + d * (pe / j);
    const double e = table[kBoysExpOffset + k] * pe;
    out[M * stride] = pref * f;
    for (int m = M - 1; m >= 0; --m) {
      f = (2.0 * T * f + e) / (2 * m + 1);
      out[m * stride] = pref * f;
    }
  } else {
    double f = 0.88622692545275801365 / sqrt(T);
    const double e = exp(-T);
    out[0] = pref * f;
    for (int m = 0; m < M; ++m) {
      f = ((2 * m + 1) * f - e) / (2.0 * T);
      out[(m + 1) * stride] = pref * f;
    }
  }
}

// Record q of the batch for primitive pairs ib of the bra and ik of the ket;
// with valid == false every field is zero (a padding slot of the batch).
template <int BM>
__device__ void setup_record(const EriArgs& a, const PairRec& bp, const PairRec& kp, int ib,
                             int ik, bool valid, int L, int P, int q, double* __restrict__ rec) {
  const int nf = rec_fields(L);
  if (!valid) {
    for (int f = 0; f < nf; ++f) rec[f * P + q] = 0.0;
    return;
  }
  const int ipb = bp.p0 + ib * bp.ps, ipk = kp.p0 + ik * kp.ps;
  const PrimRec pb = a.prims[ipb];
  const PrimRec pk = a.prims[ipk];
  const double zb = pb.zeta, zk = pk.zeta;
  const double rs = fast_rsqrt(zb + zk);
  const double ooze = rs * rs;
  const double roz = zk * ooze, roe = zb * ooze;
  // P - Q = (P - A) - (Q - C) + (A - C)
  const double PQx = pb.pax - pk.pax + (bp.A[0] - kp.A[0]);
  const double PQy = pb.pay - pk.pay + (bp.A[1] - kp.A[1]);
  const double PQz = pb.paz - pk.paz + (bp.A[2] - kp.A[2]);
  const double T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
  rec[(kPA + 0) * P + q] = pb.pax;
  rec[(kPA + 1) * P + q] = pb.pay;
  rec[(kPA + 2) * P + q] = pb.paz;
  rec[(kWP + 0) * P + q] = -roz * PQx;
  rec[(kWP + 1) * P + q] = -roz * PQy;
  rec[(kWP + 2) * P + q] = -roz * PQz;
  rec[(kQC + 0) * P + q] = pk.pax;
  rec[(kQC + 1) * P + q] = pk.pay;
  rec[(kQC + 2) * P + q] = pk.paz;
  rec[(kWQ + 0) * P + q] = roe * PQx;
  rec[(kWQ + 1) * P + q] = roe * PQy;
  rec[(kWQ + 2) * P + q] = roe * PQz;
  rec[kOo2z * P + q] = pb.oo2z;
  rec[kOo2e * P + q] = pk.oo2z;
  rec[kOo2ze * P + q] = 0.5 * ooze;
  rec[kRoz * P + q] = roz;
  rec[kRoe * P + q] = roe;
  if (a.prim_alpha) {
    const double aa = a.prim_alpha[ipb], ac = a.prim_alpha[ipk];
    rec[kWA * P + q] = 2.0 * aa;
    rec[kWB * P + q] = 2.0 * (zb - aa);
    rec[kWC * P + q] = 2.0 * ac;
  } else {
    rec[kWA * P + q] = 0.0;
    rec[kWB * P + q] = 0.0;
    rec[kWC * P + q] = 0.0;
  }
  const double pref = pb.kab * pk.kab * rs;
  if constexpr (BM == kBoysCompat) {
    for (int m = 0; m <= L; ++m)
      rec[(kG + m) * P + q] = pref * fgamma_compat(static_cast<double>(m), T);
  } else {
    boys_exact_strided(L, T, a.boys_table, pref, rec + kG * P + q, P);
  }
}

// One contracted quartet (bra pair bidx, ket pair kidx): the HRR results end in
// H = R; the outputs are read by the epilogue.
template <int BM>
__device__ void quartet(const EriArgs& a, const GenericProgram& p, int bidx, int kidx,
                        double* __restrict__ rec, double* __restrict__ R, int lane) {
  const PairRec bp = a.pairs[bidx];
  const PairRec kp = a.pairs[kidx];
  const int L = p.L;
  __syncthreads();  // the previous quartet's outputs have been read
  double* __restrict__ acc = R;
  for (int k = lane; k < p.nacc; k += 64) acc[k] = 0.0;
  const int nk = kp.np, npq = bp.np * kp.np;
  int P = 1, LP = 0;
  while (P < p.P && P < npq) {
    P <<= 1;
    ++LP;
  }
  double* __restrict__ V = R + p.nacc;
  const int q = lane & (P - 1);
  const int ostep = 64 >> LP;
  for (int b0 = 0; b0 < npq; b0 += P) {
    __syncthreads();  // the previous batch's records and slots are consumed
    if (lane < P) {
      const int j = b0 + lane;
      const bool valid = j < npq;
      setup_record<BM>(a, bp, kp, valid ? j / nk : 0, valid ? j % nk : 0, valid, L, P, lane, rec);
    }
    __syncthreads();
    for (int idx = lane; idx < (L + 2) * P; idx += 64) {
      const int s = idx >> LP, qq = idx & (P - 1);
      V[idx] = s == 0 ? 0.0 : rec[(kG + s - 1) * P + qq];
    }
    __syncthreads();
    for (int g = 0; g < p.nvgroups; ++g) {
      const uint4 gr = p.vgroups[g];
      const int side = gr.z & 1, ax = (gr.z >> 1) & 3;
      const double c0 = rec[((side ? kQC : kPA) + ax) * P + q];
      const double c1 = rec[((side ? kWQ : kWP) + ax) * P + q];
      const double o = rec[(side ? kOo2e : kOo2z) * P + q];
      const double r = rec[(side ? kRoe : kRoz) * P + q];
      const double oz = rec[kOo2ze * P + q];
      for (int op = static_cast<int>(gr.x) + (lane >> LP); op < static_cast<int>(gr.y);
           op += ostep) {
        const uint4 w = p.vops[op];
        const double n1 = static_cast<double>(w.w & 15u);
        const double n2 = static_cast<double>((w.w >> 4) & 15u);
        double v = c0 * V[(w.x >> 16) * P + q] + c1 * V[(w.y & 0xffffu) * P + q];
        v += (n1 * o) * (V[(w.y >> 16) * P + q] - r * V[(w.z & 0xffffu) * P + q]);
        v += (n2 * oz) * V[(w.z >> 16) * P + q];
        V[(w.x & 0xffffu) * P + q] = v;
      }
      if (gr.z & 8u) __syncthreads();
    }
    for (int t = lane; t < p.naccs; t += 64) {
      const uint2 e = p.accs[t];
      const double* src = V + (e.x & 0xffffu) * P;
      const int set = static_cast<int>(e.x >> 16);
      double s = 0.0;
      if (set == 0) {
        for (int qq = 0; qq < P; ++qq) s += src[qq];
      } else {
        const double* wt = rec + (kWA + set - 1) * P;
        for (int qq = 0; qq < P; ++qq) s += wt[qq] * src[qq];
      }
      acc[e.y] += s;
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

__device__ __forceinline__ double output(const GenericProgram& p, const double* __restrict__ H,
                                         int k) {
  const int4 o = p.outs[k];
  return p.out_norm[k] * (H[o.x] - static_cast<double>(o.z) * H[o.y]);
}

// 8-fold weight deg/8 of a unique quartet of pairs (bidx, kidx)
__device__ __forceinline__ double sym_weight(const PairRec& bp, const PairRec& kp, int bidx,
                                             int kidx) {
  double s = 0.125;
  if (bp.flags & 1) s *= 2.0;
  if (kp.flags & 1) s *= 2.0;
  if (bidx != kidx) s *= 2.0;
  return s;
}

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
      a.out[static_cast<size_t>(t.pad) + k] = output(p, H, k);
  } else if constexpr (MODE == kModeDiag) {
    double m = 0.0;
    for (int k = lane; k < NA * NB; k += 64) {
      const int ia = k / NB, ib = k % NB;
      m = fmax(m, fabs(output(p, H, ((ia * NB + ib) * NC + ia) * ND + ib)));
    }
    m = wave_max(m);
    if (lane == 0) a.out[kidx] = sqrt(m);
  } else if constexpr (MODE == kModeTensor) {
    for (int k = lane; k < p.nout; k += 64) {
      const double v = output(p, H, k);
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
  } else if constexpr (MODE == kModeJK) {
    // the lower-half accumulators get deg/8 of every permutation's
    // contribution; J + J^T, K + K^T restore the rest
    const double s = sym_weight(bp, kp, bidx, kidx);
    const double* __restrict__ D = a.D;
    for (int k = lane; k < p.nout; k += 64) {
      const double v = s * output(p, H, k);
      const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
      const size_t A = fa + ia, B = fb + ib, C = fc + ic, Dd = fd + id;
      unsafeAtomicAdd(&a.J[A * n + B], 2.0 * v * D[C * n + Dd]);
      unsafeAtomicAdd(&a.J[C * n + Dd], 2.0 * v * D[A * n + B]);
      unsafeAtomicAdd(&a.K[A * n + C], v * D[B * n + Dd]);
      unsafeAtomicAdd(&a.K[A * n + Dd], v * D[B * n + C]);
      unsafeAtomicAdd(&a.K[B * n + C], v * D[A * n + Dd]);
      unsafeAtomicAdd(&a.K[B * n + Dd], v * D[A * n + C]);
    }
  } else {  // kModeGrad
    // dE/dX = sum over ordered quartets of Gamma d(mu nu|la si)/dX with
    // Gamma = 2 D_mn D_ls - (D_ml D_ns + D_ms D_nl) / 2; a unique quartet
    // stands for deg = 8 s ordered ones.  Outputs: (centre A, B, C) x (x, y, z)
    // x components; centre D by translational invariance.
    const int nc = NA * NB * NC * ND;
    const double deg = 8.0 * sym_weight(bp, kp, bidx, kidx);
    const double* __restrict__ D = a.D;
    // a per-centre program holds that centre's 3 x nc outputs: map them onto
    // rows 3 centre .. 3 centre + 2 of g, the others stay zero
    const int c0 = p.centre < 0 ? 0 : 3 * p.centre, nrow = p.centre < 0 ? 9 : 3;
    double g[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) g[c] = 0.0;
    for (int k = lane; k < nc; k += 64) {
      const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
      const size_t A = fa + ia, B = fb + ib, C = fc + ic, Dd = fd + id;
      const double gam = 2.0 * D[A * n + B] * D[C * n + Dd] -
                         0.5 * (D[A * n + C] * D[B * n + Dd] + D[A * n + Dd] * D[B * n + C]);
#pragma unroll
      for (int c = 0; c < 9; ++c)
        if (c >= c0 && c < c0 + nrow) g[c] += gam * output(p, H, (c - c0) * nc + k);
    }
    // six accumulators per pair (its two atoms' x, y, z), added onto the
    // atoms once every class is done: per-atom atomics would contend
    double* __restrict__ gb = a.grad_pair + 6 * static_cast<size_t>(bidx);
    double* __restrict__ gk = a.grad_pair + 6 * static_cast<size_t>(kidx);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double gA = deg * wave_sum(g[i]);
      const double gB = deg * wave_sum(g[3 + i]);
      const double gC = deg * wave_sum(g[6 + i]);
      if (lane == 0) {
        unsafeAtomicAdd(&gb[i], gA);
        unsafeAtomicAdd(&gb[3 + i], gB);
        unsafeAtomicAdd(&gk[i], gC);
        unsafeAtomicAdd(&gk[3 + i], -(gA + gB + gC));
      }
    }
  }
}

template <int MODE, int BM>
__global__ __launch_bounds__(64) void eri_generic_kernel(const EriArgs a, const GenericProgram p) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  double* rec = lds;
  double* R = lds + p.P * rec_fields(p.L);
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
      quartet<BM>(a, p, bidx, kidx, rec, R, lane);
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
  const size_t nf = 20 + p.L + 1;
  const size_t buf = std::max(static_cast<size_t>(p.nacc) + static_cast<size_t>(p.nv) * p.P,
                              static_cast<size_t>(p.nh));
  return sizeof(double) * (static_cast<size_t>(p.P) * nf + buf);
}

// blocks: with a device tile range, the number of 4-wave blocks the per-lane
// kernels would use (this kernel launches 4 one-wave workgroups per block);
// without one, one workgroup per tile.
hipError_t eri_generic_launch(const EriArgs& a, const GenericProgram& p, int mode, int bm,
                              int blocks, hipStream_t s, size_t lds_limit) {
  unsigned grid;
  if (a.tile_range) {
    if (blocks <= 0) return hipSuccess;
    grid = static_cast<unsigned>(blocks) * 4u;
  } else {
    if (a.ntiles <= 0) return hipSuccess;
    grid = static_cast<unsigned>(a.ntiles);
  }
  if (p.P < 1 || p.P > 64 || (p.P & (p.P - 1)) != 0 || p.L > kBoysMaxM) return hipErrorInvalidValue;
  const size_t lds = eri_generic_lds_bytes(p);
  if (lds > lds_limit) return hipErrorInvalidValue;
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
    case kModeGrad:
      return launch_mode<kModeGrad, kBoysExact>(a, p, grid, lds, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace rchem

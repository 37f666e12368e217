// Contracted Cartesian-Gaussian ERIs by Head-Gordon--Pople VRR/HRR, fused with
// J/K digestion, for gfx950 (wave64).
//
// One kernel specialisation per (la lb|lc ld) class and mode.  A wavefront
// takes one Tile: a uniform bra shell pair against up to 64 ket shell pairs of
// one bucket (same primitive-pair count), one contracted quartet per lane.
//   * bra primitive-pair data are wave-uniform (scalar loads), ket data are
//     read from the transposed bucket layout (coalesced);
//   * per primitive quartet: Boys F_0..F_L, the generated VRR into contracted
//     accumulators [e0|f0] held in registers;
//   * per contracted quartet: the generated HRR, then
//       kModeJK:     digestion into J and K with 8-fold symmetry weights;
//       kModeDiag:   Schwarz bound sqrt(max |(ab|ab)|) of the lane's pair;
//       kModeTensor: all 8 permutations into a dense nbf^4 tensor (small
//                    systems; the layout of rchem's build_I, basis.rs:430-460).
// JK digestion has two forms.  With LDS digestion (the default whenever the
// bra's K rows fit), a workgroup owns a chunk of tiles of ONE bra pair (a, b):
// every K contribution lands in rows of a or b, so the block accumulates them
// in LDS (ds_add_f64) and flushes each row once, contiguously, at the end of
// the chunk; Coulomb contributions go to pair-indexed accumulators whose
// layout makes a wave's 64 updates contiguous.  Otherwise every contribution
// is a global FP64 atomic (L2).
// The reference computes the same integrals one primitive at a time with the
// THO formula (cints.c:72-149) inside JK_direct's 8-deep loop
// (basis.rs:383-428); its HGP recurrences are chgp.c:44-586.
#pragma once
#include <hip/hip_runtime.h>

#include "boys.hpp"
#include "types.hpp"

namespace rchem {

template <int LA, int LB, int LC, int LD>
struct EriClass;

#include "generated/eri_classes.inc"

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Integrals of one tile: (ab|cd) for the lane's quartet in FP64 (out), with R
// the arithmetic type of the primitive work.  With R = float the primitive
// work (Boys, VRR, HRR, contraction) runs in FP32: geometric differences are
// formed in FP64 and rounded once, and the integrals are widened to FP64
// before digestion, whose sums and atomics stay in FP64.
struct TileInfo {
  PairRec bp, kp;
  int bidx, kidx;
  bool active;
};

// The primitive-pair records a kernel of arithmetic type R reads: FP64
// records for FP64 work, the FP32 copy (half the bytes) for FP32 work.
template <typename R>
struct PrimOf {
  using T = PrimRec;
};
template <>
struct PrimOf<float> {
  using T = PrimRecF;
};
template <typename R>
__device__ __forceinline__ const typename PrimOf<R>::T* prims_of(const PrimRec* p64,
                                                                 const PrimRecF* p32) {
  if constexpr (sizeof(R) == 8) return p64;
  else return p32;
}

// Primitive loops: the ket primitive (per-lane, vector loads; with PF fetched
// one iteration ahead) is the outer loop, the bra primitive the inner one, so
// each ket record is read once per quartet.  The bra pair is wave-uniform, so
// its records come through scalar loads (the scalar cache), not LDS or vector
// memory (the kernels take the record array as a noalias argument, which is
// what lets the compiler prove the loads unclobbered by the kernel's atomics
// and issue them as scalar loads).  Records hold P - A relative to their pair's first centre; per lane
// Q - A = (Q - C) - (A - C) once per ket primitive, then P - Q = (P - A) - (Q - A)
// per primitive quartet.  `boys_tab` is the Boys grid in LDS.
template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R, bool PF = true,
          int UNR = 1, bool BF = false>
__device__ __forceinline__ void eri_compute(const EriArgs& a, const Tile t, const int lane,
                                            TileInfo& ti,
                                            double (&out)[EriClass<LA, LB, LC, LD>::NOUT],
                                            const R* __restrict__ boys_tab,
                                            const typename PrimOf<R>::T* __restrict__ prim) {
  using C = EriClass<LA, LB, LC, LD>;
  using P = typename PrimOf<R>::T;
  constexpr int L = C::L;
  static_assert(BM == kBoysExact || sizeof(R) == 8, "compat Boys is FP64 only");
  ti.active = lane < t.count;
  ti.kidx = t.ket0 + (ti.active ? lane : 0);
  ti.bidx = (MODE == kModeDiag) ? ti.kidx : t.bra;
  ti.bp = a.pairs[ti.bidx];
  ti.kp = a.pairs[ti.kidx];
  const PairRec& bp = ti.bp;
  const PairRec& kp = ti.kp;

  R acc[C::NACC];
#pragma unroll
  for (int i = 0; i < C::NACC; ++i) acc[i] = R(0);

  const R* __restrict__ table = boys_tab;
  // A - C in FP64, rounded once
  const R ACx = R(bp.A[0] - kp.A[0]), ACy = R(bp.A[1] - kp.A[1]), ACz = R(bp.A[2] - kp.A[2]);
  const P* __restrict__ kprim = prim + kp.p0;
  // the bra pair's records: wave-uniform (said so, so that the address and the
  // trip count live in scalar registers), except in the diagonal mode
  const int bp0 = MODE == kModeDiag ? bp.p0 : __builtin_amdgcn_readfirstlane(bp.p0);
  const int bps = MODE == kModeDiag ? bp.ps : __builtin_amdgcn_readfirstlane(bp.ps);
  const int bnp = MODE == kModeDiag ? bp.np : __builtin_amdgcn_readfirstlane(bp.np);
  const P* __restrict__ bprim = prim + bp0;
  // the ket pairs of a tile share their bucket, hence its primitive-pair count
  // and record stride: the ket loop is uniform too
  const int knp = MODE == kModeDiag ? kp.np : __builtin_amdgcn_readfirstlane(kp.np);
  const int kps = MODE == kModeDiag ? kp.ps : __builtin_amdgcn_readfirstlane(kp.ps);
  P nxt = kprim[0];
  for (int j = 0; j < knp; ++j) {
    // with PF the next record is in flight during this iteration's work;
    // without, the load is issued here and other waves cover its latency
    const P pk = PF ? nxt : kprim[j * kps];
    if (PF && j + 1 < knp) nxt = kprim[(j + 1) * kps];
    const R zk = R(pk.zeta), kk = R(pk.kab), oo2e = R(pk.oo2z);
    const R QCx = R(pk.pax), QCy = R(pk.pay), QCz = R(pk.paz);
    const R QAx = QCx - ACx, QAy = QCy - ACy, QAz = QCz - ACz;
    // one bra primitive pair into the accumulators
    auto prim_quartet = [&](const P& pb) __attribute__((always_inline)) {
      const R zb = R(pb.zeta);
      const R PAx = R(pb.pax), PAy = R(pb.pay), PAz = R(pb.paz);
      const R rs = fast_rsqrt(zb + zk);  // (zeta + eta)^(-1/2)
      const R ooze = rs * rs;
      const R roz = zk * ooze;  // rho / zeta
      const R roe = zb * ooze;  // rho / eta
      const R PQx = PAx - QAx, PQy = PAy - QAy, PQz = PAz - QAz;
      const R T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
      R F[L + 1];
      boys<L, BM, R, BF>(T, table, F);
      const R pref = R(pb.kab) * kk * rs;
      R G[L + 1];
#pragma unroll
      for (int m = 0; m <= L; ++m) G[m] = pref * F[m];
      C::template vrr<R>(G, PAx, PAy, PAz, -roz * PQx, -roz * PQy, -roz * PQz, QCx, QCy, QCz,
                         roe * PQx, roe * PQy, roe * PQz, R(pb.oo2z), oo2e, R(0.5) * ooze, roz,
                         roe, PQx, PQy, PQz, acc);
    };
    int i = 0;
    if constexpr (UNR == 2) {
      for (; i + 1 < bnp; i += 2) {
        const P pb0 = bprim[i * bps];
        const P pb1 = bprim[(i + 1) * bps];
        prim_quartet(pb0);
        prim_quartet(pb1);
      }
    }
    if constexpr (sizeof(R) == 4 && UNR == 1) {
      // FP32: the next bra record is loaded before this one's primitive
      // quartet is evaluated, so its scalar-load latency overlaps the work.
      // FP64 keeps the plain loop: there this form measured no faster, and
      // in the spilling L = 5, 6 per-lane kernels (quartet mode) it gave wrong
      // integrals, a cause not found; the FP32 kernels stop at L = 4.
      if (bnp > 0) {
        P cur = bprim[0];
        for (; i + 1 < bnp; ++i) {
          const P nxt_b = bprim[(i + 1) * bps];
          prim_quartet(cur);
          cur = nxt_b;
        }
        prim_quartet(cur);
      }
    } else {
      for (; i < bnp; ++i) prim_quartet(bprim[i * bps]);
    }
  }

  R outr[C::NOUT];
  C::template hrr<R>(acc, R(bp.AB[0]), R(bp.AB[1]), R(bp.AB[2]), R(kp.AB[0]), R(kp.AB[1]),
                     R(kp.AB[2]), outr);
#pragma unroll
  for (int k = 0; k < C::NOUT; ++k) out[k] = static_cast<double>(outr[k]);
}

template <typename R>
__device__ __forceinline__ const R* boys_src(const EriArgs& a) {
  if constexpr (sizeof(R) == 8) return reinterpret_cast<const R*>(a.boys_table);
  else return reinterpret_cast<const R*>(a.boys_table_f);
}

// 8-fold weight: each unique (ab|cd) stands for deg of the ordered quartets;
// the lower-half accumulators get deg/8 of each of the eight permutations'
// contributions, and J + J^T, K + K^T restore the rest.  The weight is
// 2^(e - 3), e = [a != b] + [c != d] + [ab != cd], written straight into the
// exponent field (integer operations in place of selects between doubles);
// an inactive lane gets +0.
__device__ __forceinline__ double symmetry_weight(const TileInfo& ti) {
  const int e = (ti.bp.flags & 1) + (ti.kp.flags & 1) + (ti.bidx != ti.kidx ? 1 : 0);
  const int hi = ti.active ? (1023 - 3 + e) << 20 : 0;
  return __hiloint2double(hi, 0);
}

// Element at a byte offset from a base: with a wave-uniform base (scalar
// registers) and a 32-bit per-lane offset the load is one vector instruction
// with no 64-bit address arithmetic per lane.
__device__ __forceinline__ double ld_off(const double* __restrict__ base, unsigned boff) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + boff);
}
__device__ __forceinline__ double* at_off(double* base, unsigned boff) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(base) + boff);
}

// The four exchange blocks of one quartet, each contracted with its density
// block:  K_ac += s (ab|cd) D_bd,  K_ad += s (ab|cd) D_bc,
//         K_bc += s (ab|cd) D_ad,  K_bd += s (ab|cd) D_ac.
// add(row_is_b, row_component, column_function, value) receives each sum.
// The bra functions fa, fb are wave-uniform: the density rows they index are
// uniform bases, the ket columns fc, fd 32-bit byte offsets.
template <int NA, int NB, int NC, int ND, typename Add>
__device__ __forceinline__ void exchange_blocks(const double* __restrict__ D, int n, int fa,
                                                int fb, int fc, int fd, const double* out,
                                                double s, Add add) {
  double dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
  const unsigned oc = static_cast<unsigned>(fc) * 8u, od = static_cast<unsigned>(fd) * 8u;
#pragma unroll
  for (int ib = 0; ib < NB; ++ib) {
    const double* __restrict__ row = D + static_cast<size_t>(fb + ib) * n;
#pragma unroll
    for (int id = 0; id < ND; ++id) dbd[ib * ND + id] = ld_off(row, od + 8u * id);
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) dbc[ib * NC + ic] = ld_off(row, oc + 8u * ic);
  }
#pragma unroll
  for (int ia = 0; ia < NA; ++ia) {
    const double* __restrict__ row = D + static_cast<size_t>(fa + ia) * n;
#pragma unroll
    for (int id = 0; id < ND; ++id) dad[ia * ND + id] = ld_off(row, od + 8u * id);
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) dac[ia * NC + ic] = ld_off(row, oc + 8u * ic);
  }
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) {
      double v = 0.0;
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int id = 0; id < ND; ++id) v += out[((ia * NB + ib) * NC + ic) * ND + id] * dbd[ib * ND + id];
      add(false, ia, fc + ic, s * v);
    }
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int id = 0; id < ND; ++id) {
      double v = 0.0;
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) v += out[((ia * NB + ib) * NC + ic) * ND + id] * dbc[ib * NC + ic];
      add(false, ia, fd + id, s * v);
    }
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) {
      double v = 0.0;
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int id = 0; id < ND; ++id) v += out[((ia * NB + ib) * NC + ic) * ND + id] * dad[ia * ND + id];
      add(true, ib, fc + ic, s * v);
    }
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int id = 0; id < ND; ++id) {
      double v = 0.0;
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) v += out[((ia * NB + ib) * NC + ic) * ND + id] * dac[ia * NC + ic];
      add(true, ib, fd + id, s * v);
    }
}

// Coulomb sums of one quartet: jab[ab] = sum_cd (ab|cd) D_cd and
// jcd[cd] = sum_ab (ab|cd) D_ab (unweighted).  D_ab is wave-uniform (the bra
// block).  The ket block D_cd comes from the pair-indexed copy when given
// (component k at dket_base + dket_off + 8 k dket_stride bytes), else from D.
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void coulomb_sums(const double* __restrict__ D, int n, int fa, int fb,
                                             int fc, int fd, const double* out, double* jab,
                                             double* jcd,
                                             const double* __restrict__ dket_base = nullptr,
                                             unsigned dket_off = 0, unsigned dket_stride = 0) {
  double dab[NA * NB], dcd[NC * ND];
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) dab[ia * NB + ib] = D[static_cast<size_t>(fa + ia) * n + fb + ib];
  if (dket_base) {
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) dcd[k] = ld_off(dket_base, dket_off + 8u * k * dket_stride);
  } else {
    const unsigned od = static_cast<unsigned>(fd) * 8u;
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) {
      const double* __restrict__ row = D + static_cast<size_t>(fc + ic) * n;
#pragma unroll
      for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = ld_off(row, od + 8u * id);
    }
  }
#pragma unroll
  for (int k = 0; k < NC * ND; ++k) jcd[k] = 0.0;
#pragma unroll
  for (int ab = 0; ab < NA * NB; ++ab) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) {
      const double x = out[ab * NC * ND + k];
      v += x * dcd[k];
      jcd[k] += x * dab[ab];
    }
    jab[ab] = v;
  }
}

// Tile-granular kernel body: diagonal, tensor, and JK with global atomics.
template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
__device__ __forceinline__ void eri_tile(const EriArgs& a, const Tile t, const int lane,
                                         const R* __restrict__ boys_tab,
                                         const typename PrimOf<R>::T* __restrict__ prim,
                                         const double* __restrict__ D) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  TileInfo ti;
  double out[C::NOUT];
  eri_compute<LA, LB, LC, LD, MODE, BM, R>(a, t, lane, ti, out, boys_tab, prim);
  const int fc = ti.kp.fa, fd = ti.kp.fb;
  // the bra's functions are wave-uniform except in the diagonal mode
  const int fa = MODE == kModeDiag ? ti.bp.fa : __builtin_amdgcn_readfirstlane(ti.bp.fa);
  const int fb = MODE == kModeDiag ? ti.bp.fb : __builtin_amdgcn_readfirstlane(ti.bp.fb);
  const size_t n = static_cast<size_t>(a.nbf);

  if constexpr (MODE == kModeDiag) {
    double m = 0.0;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) m = fmax(m, fabs(out[((ia * NB + ib) * NC + ia) * ND + ib]));
    if (ti.active) a.out[ti.kidx] = sqrt(m);
  } else if constexpr (MODE == kModeQuartet) {
    if (ti.active) {
#pragma unroll
      for (int k = 0; k < C::NOUT; ++k) a.out[static_cast<size_t>(t.pad) + k] = out[k];
    }
  } else if constexpr (MODE == kModeTensor) {
    if (ti.active) {
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
              const double v = out[((ia * NB + ib) * NC + ic) * ND + id];
              const size_t p = fa + ia, q = fb + ib, r = fc + ic, s = fd + id;
              double* I = a.out;
              I[((p * n + q) * n + r) * n + s] = v;
              I[((q * n + p) * n + r) * n + s] = v;
              I[((p * n + q) * n + s) * n + r] = v;
              I[((q * n + p) * n + s) * n + r] = v;
              I[((r * n + s) * n + p) * n + q] = v;
              I[((s * n + r) * n + p) * n + q] = v;
              I[((r * n + s) * n + q) * n + p] = v;
              I[((s * n + r) * n + q) * n + p] = v;
            }
    }
  } else {
    const double s = symmetry_weight(ti);
    const int ni = a.nbf;
    double jab[NA * NB], jcd[NC * ND];
    coulomb_sums<NA, NB, NC, ND>(D, ni, fa, fb, fc, fd, out, jab, jcd);
#pragma unroll
    for (int ab = 0; ab < NA * NB; ++ab) {
      const double v = wave_sum(2.0 * s * jab[ab]);
      if (lane == 0) unsafeAtomicAdd(&a.J[(fa + ab / NB) * n + fb + ab % NB], v);
    }
    if (ti.active) {
      const unsigned od = static_cast<unsigned>(fd) * 8u;
#pragma unroll
      for (int k = 0; k < NC * ND; ++k)
        unsafeAtomicAdd(at_off(a.J + (fc + k / ND) * n, od + 8u * (k % ND)), 2.0 * s * jcd[k]);
      double* K = a.K;
      exchange_blocks<NA, NB, NC, ND>(D, ni, fa, fb, fc, fd, out, s,
                                      [&](bool rb, int comp, int col, double v) {
                                        const int row = rb ? fb + comp : fa + comp;
                                        unsafeAtomicAdd(at_off(K + row * n,
                                                               static_cast<unsigned>(col) * 8u),
                                                        v);
                                      });
    }
  }
}

template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
__global__ __launch_bounds__(256) void eri_kernel(const EriArgs a,
                                                  const PrimRec* __restrict__ prims,
                                                  const PrimRecF* __restrict__ prims_f,
                                                  const double* __restrict__ D) {
  const typename PrimOf<R>::T* __restrict__ prim = prims_of<R>(prims, prims_f);
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  __shared__ R boys_tab[boys_stage_size(EriClass<LA, LB, LC, LD>::L)];
  stage_boys<EriClass<LA, LB, LC, LD>::L>(boys_tab, boys_src<R>(a));
  __syncthreads();
  // the wave index is wave-uniform; saying so lets the tile, the bra pair and
  // the bra primitives come through scalar loads
  const int w = __builtin_amdgcn_readfirstlane(wave);
  if (a.tile_range) {
    // persistent: a fixed grid drains tiles [range[0], range[1]); every wave
    // leaves the loop once its next index passes the end
    const int begin = a.tile_range[0], end = a.tile_range[1];
    for (int t = begin + w; t < end; t += gridDim.x * 4)
      eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[t], lane, boys_tab, prim, D);
  } else {
    if (w >= a.ntiles) return;  // uniform per wave
    eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[w], lane, boys_tab, prim, D);
  }
}

// LDS digestion: a fixed grid of workgroups drains chunks [range[0], range[1]).
// Dynamic LDS holds the bra's K rows: (NA + NB) x kw doubles, columns
// [kc0, kc0 + kw) (all nbf columns for one molecule).  W waves per workgroup
// share them; OCC > 0 bounds the registers so that OCC waves fit a SIMD; PF
// selects the one-ahead ket fetch, UNR the bra primitive pairs per inner
// iteration, BF the branch-free Boys function (kLdsVariants lists the shapes).
template <int LA, int LB, int LC, int LD, int BM, typename R, int W, int OCC, bool PF,
          int UNR, bool BF>
__global__ __launch_bounds__(64 * W, OCC > 0 ? OCC : 1) void eri_jk_lds_kernel(
    const EriArgs a, const PrimRec* __restrict__ prims, const PrimRecF* __restrict__ prims_f,
    const double* __restrict__ D, const double* __restrict__ Dp) {
  const typename PrimOf<R>::T* __restrict__ prim = prims_of<R>(prims, prims_f);
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  extern __shared__ double krows[];
  __shared__ R boys_tab[boys_stage_size(C::L)];
  stage_boys<C::L>(boys_tab, boys_src<R>(a));
  const size_t n = static_cast<size_t>(a.nbf);
  const int ni = a.nbf;
  const int kw = a.kw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nrow = (NA + NB) * kw;
  const int begin = a.chunk_range[0], end = a.chunk_range[1];
  for (int ch = begin + blockIdx.x; ch < end; ch += gridDim.x) {  // block-uniform
    const Chunk c = a.chunks[ch];
    const int kc0 = a.pair_kc0 ? a.pair_kc0[c.bra] : 0;
    for (int k = threadIdx.x; k < nrow; k += blockDim.x) krows[k] = 0.0;
    __syncthreads();
    double jab_acc[NA * NB];
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) jab_acc[k] = 0.0;
    for (int t = c.tb + wave; t < c.te; t += W) {  // wave-uniform
      TileInfo ti;
      double out[C::NOUT];
      eri_compute<LA, LB, LC, LD, kModeJK, BM, R, PF, UNR, BF>(a, a.tiles[t], lane, ti, out,
                                                               boys_tab, prim);
      const int fa = __builtin_amdgcn_readfirstlane(ti.bp.fa);
      const int fb = __builtin_amdgcn_readfirstlane(ti.bp.fb);
      const int fc = ti.kp.fa, fd = ti.kp.fb;
      const double s = symmetry_weight(ti);
      double jab[NA * NB], jcd[NC * ND];
      const int kq = ti.kidx - a.jp_ket.begin;
      const unsigned kq8 = static_cast<unsigned>(kq) * 8u;
      const unsigned kstride = static_cast<unsigned>(a.jp_ket.n);
      coulomb_sums<NA, NB, NC, ND>(D, ni, fa, fb, fc, fd, out, jab, jcd,
                                   Dp ? Dp + a.jp_ket.base : nullptr, kq8, kstride);
#pragma unroll
      for (int k = 0; k < NA * NB; ++k) jab_acc[k] += 2.0 * s * jab[k];
      if (ti.active) {
        double* jp_ket = a.Jp + a.jp_ket.base;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k)
          unsafeAtomicAdd(at_off(jp_ket, kq8 + 8u * k * kstride), 2.0 * s * jcd[k]);
        exchange_blocks<NA, NB, NC, ND>(D, ni, fa, fb, fc, fd, out, s,
                                        [&](bool rb, int comp, int col, double v) {
                                          atomicAdd(&krows[(rb ? NA + comp : comp) * kw + col - kc0],
                                                    v);
                                        });
      }
    }
    // bra Coulomb block: wave sums, one atomic per component per wave
    const int bq = c.bra - a.jp_bra.begin;
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) {
      const double v = wave_sum(jab_acc[k]);
      if (lane == 0 && v != 0.0) unsafeAtomicAdd(&a.Jp[a.jp_bra.base + k * a.jp_bra.n + bq], v);
    }
    __syncthreads();
    // flush the K rows of a and b, contiguous along each row
    const PairRec bp = a.pairs[c.bra];
    for (int k = threadIdx.x; k < nrow; k += blockDim.x) {
      const double v = krows[k];
      if (v != 0.0) {
        const int r = k / kw, col = kc0 + (k - r * kw);
        const int row = r < NA ? bp.fa + r : bp.fb + (r - NA);
        unsafeAtomicAdd(&a.K[row * n + col], v);
      }
    }
    __syncthreads();
  }
}

using EriLaunchFn = hipError_t (*)(const EriArgs&, int, hipStream_t);

template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
hipError_t eri_launch(const EriArgs& a, int blocks, hipStream_t s) {
  unsigned grid;
  if (a.tile_range) {
    if (blocks <= 0) return hipSuccess;
    grid = static_cast<unsigned>(blocks);
  } else {
    if (a.ntiles <= 0) return hipSuccess;
    grid = static_cast<unsigned>((a.ntiles + 3) / 4);
  }
  hipLaunchKernelGGL((eri_kernel<LA, LB, LC, LD, MODE, BM, R>), dim3(grid), dim3(256), 0, s, a,
                     a.prims, a.prims_f, a.D);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int BM, typename R, int W = 4, int OCC = 0,
          bool PF = true, int UNR = 1, bool BF = false>
hipError_t eri_jk_lds_launch(const EriArgs& a, int blocks, hipStream_t s) {
  if (blocks <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (ncart(LA) + ncart(LB)) * static_cast<size_t>(a.kw);
  hipLaunchKernelGGL((eri_jk_lds_kernel<LA, LB, LC, LD, BM, R, W, OCC, PF, UNR, BF>), dim3(blocks),
                     dim3(64 * W), lds, s, a, a.prims, a.prims_f, a.D, a.Dp);
  return hipGetLastError();
}

// The LDS-digestion launcher of shape kLdsVariants[V] (exact Boys; FP64 or FP32).
template <int LA, int LB, int LC, int LD, int V>
EriLaunchFn eri_jk_lds_variant(int bm, int fp32) {
  constexpr LdsVariant v = kLdsVariants[V];
  if (bm != kBoysExact) return nullptr;
  if (fp32) return eri_jk_lds_launch<LA, LB, LC, LD, kBoysExact, float, v.waves,
                                     v.min_waves_per_simd, v.prefetch != 0, v.unroll,
                                     v.branch_free != 0>;
  return eri_jk_lds_launch<LA, LB, LC, LD, kBoysExact, double, v.waves, v.min_waves_per_simd,
                           v.prefetch != 0, v.unroll, v.branch_free != 0>;
}

}  // namespace rchem

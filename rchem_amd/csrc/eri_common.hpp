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

// A bra primitive pair as the LDS-digestion kernel stages it, in the kernel's
// arithmetic type: PA = P - A is formed in FP64 once per chunk and rounded
// once, so the primitive loop never touches FP64 in the FP32 kernels.
template <typename R>
struct BraPrim {
  R zeta, pax, pay, paz, kab, oo2z;
};

// Primitive loops: the ket primitive (per-lane, vector loads, fetched one
// iteration ahead) is the outer loop, the bra primitive (wave-uniform) the inner
// one, so each ket record is read once per quartet.  With BRA_LDS the bra's
// primitive pairs are read from `sbra` in LDS (staged once per chunk by the
// LDS-digestion kernel) and P - Q is formed as (P - A) - (Q - A), Q - A once per
// ket primitive; otherwise they come from global memory.  `boys_tab` is the
// Boys grid in LDS.
template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R, bool BRA_LDS = false,
          bool PF = true>
__device__ __forceinline__ void eri_compute(const EriArgs& a, const Tile t, const int lane,
                                            TileInfo& ti,
                                            double (&out)[EriClass<LA, LB, LC, LD>::NOUT],
                                            const R* __restrict__ boys_tab,
                                            const BraPrim<R>* __restrict__ sbra = nullptr) {
  using C = EriClass<LA, LB, LC, LD>;
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

  const PrimRec* __restrict__ kprim = a.prims + kp.p0;
  PrimRec nxt = kprim[0];
  for (int j = 0; j < kp.np; ++j) {
    // with PF the next record is in flight during this iteration's work;
    // without, the load is issued here and other waves cover its latency
    const PrimRec pk = PF ? nxt : kprim[j * kp.ps];
    if (PF && j + 1 < kp.np) nxt = kprim[(j + 1) * kp.ps];
    const R zk = R(pk.zeta), kk = R(pk.kab), oo2e = R(pk.oo2z);
    const R QCx = R(pk.px - kp.A[0]), QCy = R(pk.py - kp.A[1]), QCz = R(pk.pz - kp.A[2]);
    if constexpr (BRA_LDS) {
      const R QAx = R(pk.px - bp.A[0]), QAy = R(pk.py - bp.A[1]), QAz = R(pk.pz - bp.A[2]);
      for (int i = 0; i < bp.np; ++i) {
        const BraPrim<R> pb = sbra[i];
        const R zb = pb.zeta;
        const R rs = fast_rsqrt(zb + zk);  // (zeta + eta)^(-1/2)
        const R ooze = rs * rs;
        const R roz = zk * ooze;  // rho / zeta
        const R roe = zb * ooze;  // rho / eta
        const R PQx = pb.pax - QAx, PQy = pb.pay - QAy, PQz = pb.paz - QAz;
        const R T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
        R F[L + 1];
        boys<L, BM, R>(T, table, F);
        const R pref = pb.kab * kk * rs;
        R G[L + 1];
#pragma unroll
        for (int m = 0; m <= L; ++m) G[m] = pref * F[m];
        C::template vrr<R>(G, pb.pax, pb.pay, pb.paz, -roz * PQx, -roz * PQy, -roz * PQz, QCx,
                           QCy, QCz, roe * PQx, roe * PQy, roe * PQz, pb.oo2z, oo2e,
                           R(0.5) * ooze, roz, roe, acc);
      }
    } else {
      for (int i = 0; i < bp.np; ++i) {
        const PrimRec pb = a.prims[bp.p0 + i * bp.ps];
        const R zb = R(pb.zeta);
        const R rs = fast_rsqrt(zb + zk);  // (zeta + eta)^(-1/2)
        const R ooze = rs * rs;
        const R roz = zk * ooze;  // rho / zeta
        const R roe = zb * ooze;  // rho / eta
        const R PQx = R(pb.px - pk.px), PQy = R(pb.py - pk.py), PQz = R(pb.pz - pk.pz);
        const R T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
        R F[L + 1];
        boys<L, BM, R>(T, table, F);
        const R pref = R(pb.kab) * kk * rs;
        R G[L + 1];
#pragma unroll
        for (int m = 0; m <= L; ++m) G[m] = pref * F[m];
        C::template vrr<R>(G, R(pb.px - bp.A[0]), R(pb.py - bp.A[1]), R(pb.pz - bp.A[2]),
                           -roz * PQx, -roz * PQy, -roz * PQz, QCx, QCy, QCz, roe * PQx,
                           roe * PQy, roe * PQz, R(pb.oo2z), oo2e, R(0.5) * ooze, roz, roe, acc);
      }
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
// contributions, and J + J^T, K + K^T restore the rest.
__device__ __forceinline__ double symmetry_weight(const TileInfo& ti) {
  double s = 0.125;
  if (ti.bp.flags & 1) s *= 2.0;
  if (ti.kp.flags & 1) s *= 2.0;
  if (ti.bidx != ti.kidx) s *= 2.0;
  return ti.active ? s : 0.0;
}

// The four exchange blocks of one quartet, each contracted with its density
// block:  K_ac += s (ab|cd) D_bd,  K_ad += s (ab|cd) D_bc,
//         K_bc += s (ab|cd) D_ad,  K_bd += s (ab|cd) D_ac.
// add(row_is_b, row_component, column_function, value) receives each sum.
template <int NA, int NB, int NC, int ND, typename Add>
__device__ __forceinline__ void exchange_blocks(const double* __restrict__ D, size_t n, int fa,
                                                int fb, int fc, int fd, const double* out,
                                                double s, Add add) {
  double dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
#pragma unroll
  for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
    for (int id = 0; id < ND; ++id) dbd[ib * ND + id] = D[(fb + ib) * n + fd + id];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) dbc[ib * NC + ic] = D[(fb + ib) * n + fc + ic];
  }
#pragma unroll
  for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
    for (int id = 0; id < ND; ++id) dad[ia * ND + id] = D[(fa + ia) * n + fd + id];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) dac[ia * NC + ic] = D[(fa + ia) * n + fc + ic];
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
// jcd[cd] = sum_ab (ab|cd) D_ab (unweighted).  The ket block D_cd comes from
// the pair-indexed copy Dp (component k at dket[k * dket_stride]) when given,
// else from D.
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void coulomb_sums(const double* __restrict__ D, size_t n, int fa,
                                             int fb, int fc, int fd, const double* out,
                                             double* jab, double* jcd,
                                             const double* __restrict__ dket = nullptr,
                                             int dket_stride = 0) {
  double dab[NA * NB], dcd[NC * ND];
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) dab[ia * NB + ib] = D[(fa + ia) * n + fb + ib];
  if (dket) {
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) dcd[k] = dket[k * dket_stride];
  } else {
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = D[(fc + ic) * n + fd + id];
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
                                         const R* __restrict__ boys_tab) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  TileInfo ti;
  double out[C::NOUT];
  eri_compute<LA, LB, LC, LD, MODE, BM, R>(a, t, lane, ti, out, boys_tab);
  const int fa = ti.bp.fa, fb = ti.bp.fb, fc = ti.kp.fa, fd = ti.kp.fb;
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
    double jab[NA * NB], jcd[NC * ND];
    coulomb_sums<NA, NB, NC, ND>(a.D, n, fa, fb, fc, fd, out, jab, jcd);
#pragma unroll
    for (int ab = 0; ab < NA * NB; ++ab) {
      const double v = wave_sum(2.0 * s * jab[ab]);
      if (lane == 0) unsafeAtomicAdd(&a.J[(fa + ab / NB) * n + fb + ab % NB], v);
    }
    if (ti.active) {
#pragma unroll
      for (int k = 0; k < NC * ND; ++k)
        unsafeAtomicAdd(&a.J[(fc + k / ND) * n + fd + k % ND], 2.0 * s * jcd[k]);
      double* __restrict__ K = a.K;
      exchange_blocks<NA, NB, NC, ND>(a.D, n, fa, fb, fc, fd, out, s,
                                      [&](bool rb, int comp, int col, double v) {
                                        const int row = rb ? fb + comp : fa + comp;
                                        unsafeAtomicAdd(&K[row * n + col], v);
                                      });
    }
  }
}

template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
__global__ __launch_bounds__(256) void eri_kernel(const EriArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  __shared__ R boys_tab[kBoysStageSize];
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
      eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[t], lane, boys_tab);
  } else {
    if (w >= a.ntiles) return;  // uniform per wave
    eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[w], lane, boys_tab);
  }
}

// LDS digestion: a fixed grid of workgroups drains chunks [range[0], range[1]).
// Dynamic LDS holds the bra's K rows: (NA + NB) x kw doubles, columns
// [kc0, kc0 + kw) (all nbf columns for one molecule).  W waves per workgroup
// share them; OCC > 0 bounds the registers so that OCC waves fit a SIMD; PF
// selects the one-ahead ket fetch (kLdsVariants lists the compiled shapes).
template <int LA, int LB, int LC, int LD, int BM, typename R, int W, int OCC, bool PF>
__global__ __launch_bounds__(64 * W, OCC > 0 ? OCC : 1) void eri_jk_lds_kernel(const EriArgs a) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  extern __shared__ double krows[];
  __shared__ BraPrim<R> sbra[kMaxPrimPairs];
  __shared__ R boys_tab[kBoysStageSize];
  stage_boys<C::L>(boys_tab, boys_src<R>(a));
  const size_t n = static_cast<size_t>(a.nbf);
  const int kw = a.kw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nrow = (NA + NB) * kw;
  const int begin = a.chunk_range[0], end = a.chunk_range[1];
  for (int ch = begin + blockIdx.x; ch < end; ch += gridDim.x) {  // block-uniform
    const Chunk c = a.chunks[ch];
    const int kc0 = a.pair_kc0 ? a.pair_kc0[c.bra] : 0;
    for (int k = threadIdx.x; k < nrow; k += blockDim.x) krows[k] = 0.0;
    {
      const PairRec bp = a.pairs[c.bra];
      for (int k = threadIdx.x; k < bp.np; k += blockDim.x) {
        const PrimRec pr = a.prims[bp.p0 + k * bp.ps];
        sbra[k] = BraPrim<R>{R(pr.zeta), R(pr.px - bp.A[0]), R(pr.py - bp.A[1]),
                             R(pr.pz - bp.A[2]), R(pr.kab), R(pr.oo2z)};
      }
    }
    __syncthreads();
    double jab_acc[NA * NB];
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) jab_acc[k] = 0.0;
    for (int t = c.tb + wave; t < c.te; t += W) {  // wave-uniform
      TileInfo ti;
      double out[C::NOUT];
      eri_compute<LA, LB, LC, LD, kModeJK, BM, R, true, PF>(a, a.tiles[t], lane, ti, out,
                                                            boys_tab, sbra);
      const int fa = ti.bp.fa, fb = ti.bp.fb, fc = ti.kp.fa, fd = ti.kp.fb;
      const double s = symmetry_weight(ti);
      double jab[NA * NB], jcd[NC * ND];
      const int kq = ti.kidx - a.jp_ket.begin;
      coulomb_sums<NA, NB, NC, ND>(a.D, n, fa, fb, fc, fd, out, jab, jcd,
                                   a.Dp ? a.Dp + a.jp_ket.base + kq : nullptr, a.jp_ket.n);
#pragma unroll
      for (int k = 0; k < NA * NB; ++k) jab_acc[k] += 2.0 * s * jab[k];
      if (ti.active) {
#pragma unroll
        for (int k = 0; k < NC * ND; ++k)
          unsafeAtomicAdd(&a.Jp[a.jp_ket.base + k * a.jp_ket.n + kq], 2.0 * s * jcd[k]);
        exchange_blocks<NA, NB, NC, ND>(a.D, n, fa, fb, fc, fd, out, s,
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
  hipLaunchKernelGGL((eri_kernel<LA, LB, LC, LD, MODE, BM, R>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int BM, typename R, int W = 4, int OCC = 0,
          bool PF = true>
hipError_t eri_jk_lds_launch(const EriArgs& a, int blocks, hipStream_t s) {
  if (blocks <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (ncart(LA) + ncart(LB)) * static_cast<size_t>(a.kw);
  hipLaunchKernelGGL((eri_jk_lds_kernel<LA, LB, LC, LD, BM, R, W, OCC, PF>), dim3(blocks),
                     dim3(64 * W), lds, s, a);
  return hipGetLastError();
}

// The LDS-digestion launcher of shape kLdsVariants[V] (exact Boys; FP64 or FP32).
template <int LA, int LB, int LC, int LD, int V>
EriLaunchFn eri_jk_lds_variant(int bm, int fp32) {
  constexpr LdsVariant v = kLdsVariants[V];
  if (bm != kBoysExact) return nullptr;
  if (fp32) return eri_jk_lds_launch<LA, LB, LC, LD, kBoysExact, float, v.waves,
                                     v.min_waves_per_simd, v.prefetch != 0>;
  return eri_jk_lds_launch<LA, LB, LC, LD, kBoysExact, double, v.waves, v.min_waves_per_simd,
                           v.prefetch != 0>;
}

}  // namespace rchem

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
//       kModeJK:     digestion into lower-half J and K with 8-fold symmetry
//                    weights; the bra-row J block is reduced across the wave
//                    before a single atomic, the rest are FP64 atomics in L2;
//       kModeDiag:   Schwarz bound sqrt(max |(ab|ab)|) of the lane's pair;
//       kModeTensor: all 8 permutations into a dense nbf^4 tensor (small
//                    systems; the layout of rchem's build_I, basis.rs:430-460).
// The reference computes the same integrals one primitive at a time with the
// THO formula (cints.c:72-149) inside JK_direct's 8-deep loop
// (basis.rs:383-428); its HGP recurrences are chgp.c:44-586.
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

// One tile (one bra pair x up to 64 ket pairs) in arithmetic type R.  With
// R = float the primitive work (Boys, VRR, HRR, contraction) runs in FP32:
// geometric differences are formed in FP64 and rounded once, and the
// integrals are widened to FP64 before digestion, whose sums and atomics stay
// in FP64.
template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
__device__ __forceinline__ void eri_tile(const EriArgs& a, const Tile t, const int lane) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int L = C::L;
  static_assert(BM == kBoysExact || sizeof(R) == 8, "compat Boys is FP64 only");
  const bool active = lane < t.count;
  const int kidx = t.ket0 + (active ? lane : 0);
  const int bidx = (MODE == kModeDiag) ? kidx : t.bra;

  const PairRec bp = a.pairs[bidx];
  const PairRec kp = a.pairs[kidx];

  R acc[C::NACC];
#pragma unroll
  for (int i = 0; i < C::NACC; ++i) acc[i] = R(0);

  const R* __restrict__ table;
  if constexpr (sizeof(R) == 8) table = reinterpret_cast<const R*>(a.boys_table);
  else table = reinterpret_cast<const R*>(a.boys_table_f);

  for (int i = 0; i < bp.np; ++i) {
    const PrimRec pb = a.prims[bp.p0 + i * bp.ps];
    const R PAx = R(pb.px - bp.A[0]), PAy = R(pb.py - bp.A[1]), PAz = R(pb.pz - bp.A[2]);
    const R zb = R(pb.zeta), kb = R(pb.kab), oo2z = R(pb.oo2z);
    for (int j = 0; j < kp.np; ++j) {
      const PrimRec pk = a.prims[kp.p0 + j * kp.ps];
      const R zk = R(pk.zeta);
      const R ooze = R(1) / (zb + zk);
      const R roz = zk * ooze;  // rho / zeta
      const R roe = zb * ooze;  // rho / eta
      const R PQx = R(pb.px - pk.px), PQy = R(pb.py - pk.py), PQz = R(pb.pz - pk.pz);
      const R T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
      R F[L + 1];
      boys<L, BM, R>(T, table, F);
      const R pref = kb * R(pk.kab) * sqrt(ooze);
      R G[L + 1];
#pragma unroll
      for (int m = 0; m <= L; ++m) G[m] = pref * F[m];
      C::template vrr<R>(G, PAx, PAy, PAz, -roz * PQx, -roz * PQy, -roz * PQz,
                         R(pk.px - kp.A[0]), R(pk.py - kp.A[1]), R(pk.pz - kp.A[2]), roe * PQx,
                         roe * PQy, roe * PQz, oo2z, R(pk.oo2z), R(0.5) * ooze, roz, roe, acc);
    }
  }

  R outr[C::NOUT];
  C::template hrr<R>(acc, R(bp.AB[0]), R(bp.AB[1]), R(bp.AB[2]), R(kp.AB[0]), R(kp.AB[1]),
                     R(kp.AB[2]), outr);
  double out[C::NOUT];
#pragma unroll
  for (int k = 0; k < C::NOUT; ++k) out[k] = static_cast<double>(outr[k]);

  const int fa = bp.fa, fb = bp.fb, fc = kp.fa, fd = kp.fb;
  const size_t n = static_cast<size_t>(a.nbf);

  if constexpr (MODE == kModeDiag) {
    double m = 0.0;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) m = fmax(m, fabs(out[((ia * NB + ib) * NC + ia) * ND + ib]));
    if (active) a.out[kidx] = sqrt(m);
  } else if constexpr (MODE == kModeTensor) {
    if (active) {
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
    // 8-fold weight: each unique (ab|cd) stands for deg of the ordered
    // quartets; the lower-half accumulators get deg/8 of each of the eight
    // permutations' contributions, and J + J^T, K + K^T restore the rest.
    double s = 0.125;
    if (bp.flags & 1) s *= 2.0;
    if (kp.flags & 1) s *= 2.0;
    if (bidx != kidx) s *= 2.0;
    if (!active) s = 0.0;
    const double* __restrict__ D = a.D;

    double dab[NA * NB], dcd[NC * ND];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) dab[ia * NB + ib] = D[(fa + ia) * n + fb + ib];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = D[(fc + ic) * n + fd + id];

    // Coulomb: J_ab += 2s sum_cd (ab|cd) D_cd ; J_cd += 2s sum_ab (ab|cd) D_ab
    double jcd[NC * ND];
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) jcd[k] = 0.0;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        double jab = 0.0;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k) {
          const double v = out[(ia * NB + ib) * NC * ND + k];
          jab += v * dcd[k];
          jcd[k] += v * dab[ia * NB + ib];
        }
        jab = wave_sum(2.0 * s * jab);
        if (lane == 0) unsafeAtomicAdd(&a.J[(fa + ia) * n + fb + ib], jab);
      }
    if (active) {
#pragma unroll
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id)
          unsafeAtomicAdd(&a.J[(fc + ic) * n + fd + id], 2.0 * s * jcd[ic * ND + id]);
    }

    // Exchange: K_ac += s (ab|cd) D_bd, K_ad += s (ab|cd) D_bc,
    //           K_bc += s (ab|cd) D_ad, K_bd += s (ab|cd) D_ac.
    if (active) {
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) {
          double kac = 0.0;
#pragma unroll
          for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id)
              kac += out[((ia * NB + ib) * NC + ic) * ND + id] * D[(fb + ib) * n + fd + id];
          unsafeAtomicAdd(&a.K[(fa + ia) * n + fc + ic], s * kac);
        }
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          double kad = 0.0;
#pragma unroll
          for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
              kad += out[((ia * NB + ib) * NC + ic) * ND + id] * D[(fb + ib) * n + fc + ic];
          unsafeAtomicAdd(&a.K[(fa + ia) * n + fd + id], s * kad);
        }
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) {
          double kbc = 0.0;
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int id = 0; id < ND; ++id)
              kbc += out[((ia * NB + ib) * NC + ic) * ND + id] * D[(fa + ia) * n + fd + id];
          unsafeAtomicAdd(&a.K[(fb + ib) * n + fc + ic], s * kbc);
        }
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          double kbd = 0.0;
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
              kbd += out[((ia * NB + ib) * NC + ic) * ND + id] * D[(fa + ia) * n + fc + ic];
          unsafeAtomicAdd(&a.K[(fb + ib) * n + fd + id], s * kbd);
        }
    }
  }
}


template <int LA, int LB, int LC, int LD, int MODE, int BM, typename R>
__global__ __launch_bounds__(256) void eri_kernel(const EriArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a.tile_range) {
    // persistent: a fixed grid drains tiles [range[0], range[1]); every wave
    // leaves the loop once its next index passes the end
    const int begin = a.tile_range[0], end = a.tile_range[1];
    for (int t = begin + wave; t < end; t += gridDim.x * 4)
      eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[t], lane);
  } else {
    if (wave >= a.ntiles) return;  // uniform per wave
    eri_tile<LA, LB, LC, LD, MODE, BM, R>(a, a.tiles[wave], lane);
  }
}

// J = Jh + Jh^T, K = Kh + Kh^T (the lower-half accumulators of eri_kernel).
__global__ void symmetrize_kernel(const double* __restrict__ Jh, const double* __restrict__ Kh,
                                  double* __restrict__ J, double* __restrict__ K, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const size_t ij = static_cast<size_t>(i) * n + j, ji = static_cast<size_t>(j) * n + i;
  J[ij] = Jh[ij] + Jh[ji];
  K[ij] = Kh[ij] + Kh[ji];
}

using EriLaunchFn = hipError_t (*)(const EriArgs&, int, hipStream_t);

// grid: ntiles/4 blocks for a host-sized launch, `blocks` for a persistent one
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

// Launcher lookup for (class, mode, Boys mode, precision); nullptr when that
// combination is not compiled into this library.  fp32 = 1 selects the FP32
// arithmetic path (JK mode, exact Boys only).
EriLaunchFn eri_launcher(int la, int lb, int lc, int ld, int mode, int bm, int fp32) {
#define RCHEM_ERI_CASE(A, B, C, D)                                                             \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    if (fp32) {                                                                               \
      if (mode == kModeJK && bm == kBoysExact) return eri_launch<A, B, C, D, kModeJK, kBoysExact, float>; \
      return nullptr;                                                                         \
    }                                                                                         \
    if (bm == kBoysExact) {                                                                   \
      if (mode == kModeJK) return eri_launch<A, B, C, D, kModeJK, kBoysExact, double>;        \
      if (mode == kModeDiag) return eri_launch<A, B, C, D, kModeDiag, kBoysExact, double>;    \
      if (mode == kModeTensor) return eri_launch<A, B, C, D, kModeTensor, kBoysExact, double>; \
    } else {                                                                                  \
      if (mode == kModeJK) return eri_launch<A, B, C, D, kModeJK, kBoysCompat, double>;       \
      if (mode == kModeDiag) return eri_launch<A, B, C, D, kModeDiag, kBoysCompat, double>;   \
      if (mode == kModeTensor) return eri_launch<A, B, C, D, kModeTensor, kBoysCompat, double>; \
    }                                                                                         \
    return nullptr;                                                                           \
  }
  RCHEM_ERI_CLASSES_SP(RCHEM_ERI_CASE)
#undef RCHEM_ERI_CASE
  return nullptr;
}

hipError_t symmetrize_launch(const double* Jh, const double* Kh, double* J, double* K, int n,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  dim3 grid((n + 255) / 256, n);
  hipLaunchKernelGGL(symmetrize_kernel, grid, dim3(256), 0, s, Jh, Kh, J, K, n);
  return hipGetLastError();
}

}  // namespace rchem

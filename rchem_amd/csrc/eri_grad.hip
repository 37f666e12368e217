// Nuclear gradient of the two-electron energy, one contracted quartet per
// lane, for the classes whose derivative recurrences fit one lane's registers
// (generated/eri_grad_classes.inc, from gen_eri_grad.py); the other classes
// run on the wave-cooperative kernel (eri_generic.hip) with the same tables.
//
// A wavefront takes one Tile (one bra pair against up to 64 ket pairs).  Per
// primitive quartet: Boys F_0..F_{L+1}, the VRR, accumulation into the four
// contracted sets (weights 1, 2 alpha_a, 2 alpha_b, 2 alpha_c); per quartet:
// the HRR to d(ab|cd)/dA, /dB, /dC, contracted on the spot with
//   Gamma = 2 D_ab D_cd - (D_ac D_bd + D_ad D_bc) / 2
// and the quartet's degeneracy deg = 8 s (s the 8-fold symmetry weight of the
// Fock build), so dE/dX = sum deg Gamma d(ab|cd)/dX.  Centre D follows from
// translational invariance, d/dD = -(d/dA + d/dB + d/dC).  Contributions go
// to six accumulators per shell pair (its two atoms' x, y, z): the bra pair's
// once per wave, each ket pair's per lane -- thousands of addresses, where
// per-atom accumulators would have every lane of the chip contend for a few
// hundred -- and grad_pairs_kernel adds them onto the atoms at the end.
#include <hip/hip_runtime.h>

#include "boys.hpp"
#include "types.hpp"

namespace rchem {

template <int LA, int LB, int LC, int LD>
struct EriGrad;

#include "generated/eri_grad_classes.inc"

namespace {

__host__ __device__ constexpr int nc1(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ double wave_sum_g(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void eri_grad_kernel(const EriArgs a) {
  using G = EriGrad<LA, LB, LC, LD>;
  constexpr int L = G::L;
  constexpr int NA = nc1(LA), NB = nc1(LB), NC = nc1(LC), ND = nc1(LD);
  constexpr int NQ = NA * NB * NC * ND;
  static_assert(G::NOUT == 9 * NQ, "gradient outputs: 3 centres x 3 axes x components");
  __shared__ double boys_tab[kBoysStageSize];
  stage_boys<L>(boys_tab, a.boys_table);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (w >= a.ntiles) return;  // wave-uniform
  const Tile t = a.tiles[w];
  const bool active = lane < t.count;
  const int kidx = t.ket0 + (active ? lane : 0);
  const PairRec bp = a.pairs[t.bra];
  const PairRec kp = a.pairs[kidx];

  double acc[G::NACC];
#pragma unroll
  for (int i = 0; i < G::NACC; ++i) acc[i] = 0.0;
  for (int j = 0; j < kp.np; ++j) {
    const int ik = kp.p0 + j * kp.ps;
    const PrimRec pk = a.prims[ik];
    const double wC = 2.0 * a.prim_alpha[ik];
    const double QCx = pk.pax, QCy = pk.pay, QCz = pk.paz;
    // Q - A = (Q - C) - (A - C)
    const double QAx = QCx - (bp.A[0] - kp.A[0]), QAy = QCy - (bp.A[1] - kp.A[1]),
                 QAz = QCz - (bp.A[2] - kp.A[2]);
    for (int i = 0; i < bp.np; ++i) {
      const int ib = bp.p0 + i * bp.ps;
      const PrimRec pb = a.prims[ib];
      const double aa = a.prim_alpha[ib];
      const double zb = pb.zeta, zk = pk.zeta;
      const double rs = fast_rsqrt(zb + zk);
      const double ooze = rs * rs;
      const double roz = zk * ooze, roe = zb * ooze;
      const double PQx = pb.pax - QAx, PQy = pb.pay - QAy, PQz = pb.paz - QAz;
      const double T = zb * roz * (PQx * PQx + PQy * PQy + PQz * PQz);
      double F[L + 1];
      boys_exact<L, double>(T, boys_tab, F);
      const double pref = pb.kab * pk.kab * rs;
      double Gm[L + 1];
#pragma unroll
      for (int m = 0; m <= L; ++m) Gm[m] = pref * F[m];
      G::template vrr<double>(Gm, pb.pax, pb.pay, pb.paz, -roz * PQx,
                              -roz * PQy, -roz * PQz, QCx, QCy, QCz, roe * PQx, roe * PQy,
                              roe * PQz, pb.oo2z, pk.oo2z, 0.5 * ooze, roz, roe, 2.0 * aa,
                              2.0 * (zb - aa), wC, acc);
    }
  }
  double out[G::NOUT];
  G::template hrr<double>(acc, bp.AB[0], bp.AB[1], bp.AB[2], kp.AB[0], kp.AB[1], kp.AB[2], out);

  const size_t n = static_cast<size_t>(a.nbf);
  const double* __restrict__ D = a.D;
  const int fa = bp.fa, fb = bp.fb, fc = kp.fa, fd = kp.fb;
  double g[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) g[c] = 0.0;
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          const size_t A = fa + ia, B = fb + ib, C = fc + ic, Dd = fd + id;
          const double gam = 2.0 * D[A * n + B] * D[C * n + Dd] -
                             0.5 * (D[A * n + C] * D[B * n + Dd] + D[A * n + Dd] * D[B * n + C]);
          const int k = ((ia * NB + ib) * NC + ic) * ND + id;
#pragma unroll
          for (int c = 0; c < 9; ++c) g[c] += gam * out[c * NQ + k];
        }
  // degeneracy: ordered quartets this unique one stands for
  double deg = 0.0;
  if (active) {
    deg = 1.0;
    if (bp.flags & 1) deg *= 2.0;
    if (kp.flags & 1) deg *= 2.0;
    if (t.bra != kidx) deg *= 2.0;
  }
  double* __restrict__ gb = a.grad_pair + 6 * static_cast<size_t>(t.bra);
  double* __restrict__ gk = a.grad_pair + 6 * static_cast<size_t>(kidx);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double gA = deg * g[i], gB = deg * g[3 + i], gC = deg * g[6 + i];
    const double sA = wave_sum_g(gA), sB = wave_sum_g(gB);
    if (lane == 0) {
      unsafeAtomicAdd(&gb[i], sA);
      unsafeAtomicAdd(&gb[3 + i], sB);
    }
    if (active) {
      unsafeAtomicAdd(&gk[i], gC);
      unsafeAtomicAdd(&gk[3 + i], -(gA + gB + gC));
    }
  }
}

// grad[atom] += the pair accumulators of every pair on that atom
__global__ void grad_pairs_kernel(const double* __restrict__ gp, const int2* __restrict__ atoms,
                                  int npairs, double* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int2 at = atoms[p];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double x = gp[6 * static_cast<size_t>(p) + i], y = gp[6 * static_cast<size_t>(p) + 3 + i];
    if (x != 0.0) unsafeAtomicAdd(&grad[3 * at.x + i], x);
    if (y != 0.0) unsafeAtomicAdd(&grad[3 * at.y + i], y);
  }
}

template <int LA, int LB, int LC, int LD>
hipError_t eri_grad_launch_t(const EriArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((eri_grad_kernel<LA, LB, LC, LD>), dim3((a.ntiles + 3) / 4), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

}  // namespace

// Adds the pair accumulators (6 per pair) onto the atoms.
hipError_t grad_pairs_launch(const double* gp, const int2* atoms, int npairs, double* grad,
                             hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(grad_pairs_kernel, dim3((npairs + 255) / 256), dim3(256), 0, s, gp, atoms,
                     npairs, grad);
  return hipGetLastError();
}

// The per-lane gradient launcher of class (la lb|lc ld), or nullptr when the
// class runs on the wave-cooperative kernel.  Uses a.tiles/a.ntiles (host
// tiles), a.D, a.prim_alpha, a.grad_pair (zeroed by the caller).
using EriGradFn = hipError_t (*)(const EriArgs&, hipStream_t);
EriGradFn eri_grad_launcher(int la, int lb, int lc, int ld) {
#define RCHEM_GRAD_CASE(A, B, C, D) \
  if (la == A && lb == B && lc == C && ld == D) return eri_grad_launch_t<A, B, C, D>;
  RCHEM_ERI_GRAD_CLASSES(RCHEM_GRAD_CASE)
#undef RCHEM_GRAD_CASE
  return nullptr;
}

}  // namespace rchem

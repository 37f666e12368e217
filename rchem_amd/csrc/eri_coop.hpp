// Cooperative parts of the ERI kernels in which one wavefront evaluates one
// contracted shell quartet (the wave-cooperative kernel, eri_generic.hip, and
// the streaming kernel, eri_stream.hpp): the horizontal recurrence on the
// contracted [e0|f0] accumulators from a program's operation tables, and the
// epilogue that takes the outputs norm (H[plus] - n H[minus]) for each mode.
// The HRR is chgp.c:44-135's recurrence (tables from gen_eri_prog.py).
#pragma once
#include <hip/hip_runtime.h>

#include "types.hpp"

namespace rchem {

__device__ __forceinline__ double coop_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double coop_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Barrier of one wave's LDS traffic: with W = 1 the workgroup barrier; with
// several independent waves per workgroup a wave-level one (the compiler may
// not move LDS accesses across it, and the wave's own LDS operations, which
// the LDS unit executes in order, have returned).
template <int W>
__device__ __forceinline__ void wsync() {
  if constexpr (W == 1) {
    __syncthreads();
  } else {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt, expcnt untouched
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
}

// The HRR and output tables of a program, in device memory or staged in LDS.
struct CoopTables {
  const uint4* hgroups;
  const uint2* hops;
  const int4* outs;
  const double* out_norm;
};

// Bytes of the tables staged in LDS (16-byte aligned pieces).
__host__ __device__ inline size_t coop_table_bytes(const GenericProgram& p) {
  auto al = [](size_t b) { return (b + 15) & ~static_cast<size_t>(15); };
  return al(16 * static_cast<size_t>(p.nhgroups)) + al(8 * static_cast<size_t>(p.nhops)) +
         al(16 * static_cast<size_t>(p.nout)) + al(8 * static_cast<size_t>(p.nout));
}

// The program's tables in device memory, or copied by every thread of the
// workgroup to LDS at dst (coop_table_bytes(p) bytes; a barrier must follow).
__device__ inline CoopTables coop_tables(const GenericProgram& p, char* dst, bool stage) {
  if (!stage) return CoopTables{p.hgroups, p.hops, p.outs, p.out_norm};
  auto al = [](size_t b) { return (b + 15) & ~static_cast<size_t>(15); };
  const int nt = static_cast<int>(blockDim.x);
  uint4* hg = reinterpret_cast<uint4*>(dst);
  uint2* ho = reinterpret_cast<uint2*>(dst + al(16 * static_cast<size_t>(p.nhgroups)));
  int4* ou = reinterpret_cast<int4*>(reinterpret_cast<char*>(ho) + al(8 * static_cast<size_t>(p.nhops)));
  double* on = reinterpret_cast<double*>(reinterpret_cast<char*>(ou) +
                                         al(16 * static_cast<size_t>(p.nout)));
  for (int i = threadIdx.x; i < p.nhgroups; i += nt) hg[i] = p.hgroups[i];
  for (int i = threadIdx.x; i < p.nhops; i += nt) ho[i] = p.hops[i];
  for (int i = threadIdx.x; i < p.nout; i += nt) {
    ou[i] = p.outs[i];
    on[i] = p.out_norm[i];
  }
  return CoopTables{hg, ho, ou, on};
}

// Horizontal recurrence in place on H (the first nacc slots hold the
// contracted accumulators, the program's layout) with AB = A - B of the bra
// pair and CD = C - D of the ket pair: groups of independent operations, a
// wave barrier at each level's end.
template <int W>
__device__ __forceinline__ void coop_hrr(const CoopTables& tb, int nhgroups,
                                         double* __restrict__ H, double ABx, double ABy,
                                         double ABz, double CDx, double CDy, double CDz,
                                         int lane) {
  for (int g = 0; g < nhgroups; ++g) {
    const uint4 gr = tb.hgroups[g];
    const int side = gr.z & 1, ax = (gr.z >> 1) & 3;
    const double xb = ax == 0 ? ABx : (ax == 1 ? ABy : ABz);
    const double xk = ax == 0 ? CDx : (ax == 1 ? CDy : CDz);
    const double X = side ? xk : xb;
    // independent within a group: four operation words in flight at once
    for (int op0 = static_cast<int>(gr.x) + lane; op0 < static_cast<int>(gr.y); op0 += 256) {
      uint2 w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int op = op0 + 64 * u;
        w[u] = tb.hops[op < static_cast<int>(gr.y) ? op : op0];
      }
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = H[w[u].x >> 16] + X * H[w[u].y];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (op0 + 64 * u < static_cast<int>(gr.y)) H[w[u].x & 0xffffu] = v[u];
    }
    if (gr.z & 8u) wsync<W>();
  }
  wsync<W>();
}

// Density elements a quartet's digestion and gradient read: the six blocks
// D_ab, D_cd, D_bd, D_bc, D_ad, D_ac (at most 216 doubles), gathered into
// LDS by the whole wave in one pass (one memory latency instead of one per
// output).
constexpr int kCoopDBlock = 216;
struct DBlocks {
  int ab, cd, bd, bc, ad, ac;  // offsets into the gathered array
};
__device__ __forceinline__ DBlocks gather_dblocks(const double* __restrict__ D, size_t n, int fa,
                                                  int fb, int fc, int fd, int NA, int NB, int NC,
                                                  int ND, double* __restrict__ dst, int lane) {
  DBlocks o;
  o.ab = 0;
  o.cd = o.ab + NA * NB;
  o.bd = o.cd + NC * ND;
  o.bc = o.bd + NB * ND;
  o.ad = o.bc + NB * NC;
  o.ac = o.ad + NA * ND;
  const int total = o.ac + NA * NC;
  for (int k = lane; k < total; k += 64) {
    int r0, c0, nc, i;
    if (k < o.cd) { r0 = fa; c0 = fb; nc = NB; i = k - o.ab; }
    else if (k < o.bd) { r0 = fc; c0 = fd; nc = ND; i = k - o.cd; }
    else if (k < o.bc) { r0 = fb; c0 = fd; nc = ND; i = k - o.bd; }
    else if (k < o.ad) { r0 = fb; c0 = fc; nc = NC; i = k - o.bc; }
    else if (k < o.ac) { r0 = fa; c0 = fd; nc = ND; i = k - o.ad; }
    else { r0 = fa; c0 = fc; nc = NC; i = k - o.ac; }
    dst[k] = D[static_cast<size_t>(r0 + i / nc) * n + c0 + i % nc];
  }
  return o;
}

__device__ __forceinline__ int ncart_rt(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ double output(const CoopTables& tb, const double* __restrict__ H,
                                         int k) {
  const int4 o = tb.outs[k];
  return tb.out_norm[k] * (H[o.x] - static_cast<double>(o.z) * H[o.y]);
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

// Outputs of one quartet for the mode.  dl: kCoopDBlock doubles of this
// wave's LDS for the gathered density blocks (J/K and gradient modes).
template <int W, int MODE>
__device__ void epilogue(const EriArgs& a, const GenericProgram& p, const CoopTables& tb,
                         const Tile& t, int bidx, int kidx, const PairRec& bp,
                         const PairRec& kp, const double* __restrict__ H,
                         double* __restrict__ dl, int lane) {
  const int NA = ncart_rt(p.la), NB = ncart_rt(p.lb), NC = ncart_rt(p.lc), ND = ncart_rt(p.ld);
  const size_t n = static_cast<size_t>(a.nbf);
  const int fa = bp.fa, fb = bp.fb, fc = kp.fa, fd = kp.fb;
  if constexpr (MODE == kModeQuartet) {
    for (int k = lane; k < p.nout; k += 64)
      a.out[static_cast<size_t>(t.pad) + k] = output(tb, H, k);
  } else if constexpr (MODE == kModeDiag) {
    double m = 0.0;
    for (int k = lane; k < NA * NB; k += 64) {
      const int ia = k / NB, ib = k % NB;
      m = fmax(m, fabs(output(tb, H, ((ia * NB + ib) * NC + ia) * ND + ib)));
    }
    m = coop_wave_max(m);
    if (lane == 0) a.out[kidx] = sqrt(m);
  } else if constexpr (MODE == kModeTensor) {
    for (int k = lane; k < p.nout; k += 64) {
      const double v = output(tb, H, k);
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
    // contribution; J + J^T, K + K^T restore the rest.  The quartet's
    // integrals are contracted with the gathered density blocks first, so a
    // quartet adds NA NB + NC ND + (NA + NB)(NC + ND) values to J and K
    // instead of six per integral (fewer atomics in flight for the wave to
    // wait on before its next quartet's loads).
    const double s = sym_weight(bp, kp, bidx, kidx);
    const DBlocks o = gather_dblocks(a.D, n, fa, fb, fc, fd, NA, NB, NC, ND, dl, lane);
    wsync<W>();
    const int nab = NA * NB, ncd = NC * ND, nac = NA * NC, nad = NA * ND, nbc = NB * NC;
    const int nres = nab + ncd + nac + nad + nbc + NB * ND;
    auto O = [&](int ia, int ib, int ic, int id) {
      return output(tb, H, ((ia * NB + ib) * NC + ic) * ND + id);
    };
    for (int r = lane; r < nres; r += 64) {
      double v = 0.0;
      double* dst;
      if (r < nab) {  // J_ab += 2 s sum_cd (ab|cd) D_cd
        const int ia = r / NB, ib = r % NB;
        for (int ic = 0; ic < NC; ++ic)
          for (int id = 0; id < ND; ++id) v += O(ia, ib, ic, id) * dl[o.cd + ic * ND + id];
        v *= 2.0;
        dst = &a.J[(fa + ia) * n + fb + ib];
      } else if (r < nab + ncd) {  // J_cd += 2 s sum_ab (ab|cd) D_ab
        const int k = r - nab, ic = k / ND, id = k % ND;
        for (int ia = 0; ia < NA; ++ia)
          for (int ib = 0; ib < NB; ++ib) v += O(ia, ib, ic, id) * dl[o.ab + ia * NB + ib];
        v *= 2.0;
        dst = &a.J[(fc + ic) * n + fd + id];
      } else if (r < nab + ncd + nac) {  // K_ac += s sum_bd (ab|cd) D_bd
        const int k = r - nab - ncd, ia = k / NC, ic = k % NC;
        for (int ib = 0; ib < NB; ++ib)
          for (int id = 0; id < ND; ++id) v += O(ia, ib, ic, id) * dl[o.bd + ib * ND + id];
        dst = &a.K[(fa + ia) * n + fc + ic];
      } else if (r < nab + ncd + nac + nad) {  // K_ad += s sum_bc (ab|cd) D_bc
        const int k = r - nab - ncd - nac, ia = k / ND, id = k % ND;
        for (int ib = 0; ib < NB; ++ib)
          for (int ic = 0; ic < NC; ++ic) v += O(ia, ib, ic, id) * dl[o.bc + ib * NC + ic];
        dst = &a.K[(fa + ia) * n + fd + id];
      } else if (r < nab + ncd + nac + nad + nbc) {  // K_bc += s sum_ad (ab|cd) D_ad
        const int k = r - nab - ncd - nac - nad, ib = k / NC, ic = k % NC;
        for (int ia = 0; ia < NA; ++ia)
          for (int id = 0; id < ND; ++id) v += O(ia, ib, ic, id) * dl[o.ad + ia * ND + id];
        dst = &a.K[(fb + ib) * n + fc + ic];
      } else {  // K_bd += s sum_ac (ab|cd) D_ac
        const int k = r - nab - ncd - nac - nad - nbc, ib = k / ND, id = k % ND;
        for (int ia = 0; ia < NA; ++ia)
          for (int ic = 0; ic < NC; ++ic) v += O(ia, ib, ic, id) * dl[o.ac + ia * NC + ic];
        dst = &a.K[(fb + ib) * n + fd + id];
      }
      unsafeAtomicAdd(dst, s * v);
    }
  } else {  // kModeGrad
    // dE/dX = sum over ordered quartets of Gamma d(mu nu|la si)/dX with
    // Gamma = 2 D_mn D_ls - (D_ml D_ns + D_ms D_nl) / 2; a unique quartet
    // stands for deg = 8 s ordered ones.  Outputs: (centre A, B, C) x (x, y, z)
    // x components; centre D by translational invariance.
    const int nc = NA * NB * NC * ND;
    const double deg = 8.0 * sym_weight(bp, kp, bidx, kidx);
    const DBlocks o = gather_dblocks(a.D, n, fa, fb, fc, fd, NA, NB, NC, ND, dl, lane);
    wsync<W>();
    // a per-centre program holds that centre's 3 x nc outputs: map them onto
    // rows 3 centre .. 3 centre + 2 of g, the others stay zero
    const int c0 = p.centre < 0 ? 0 : 3 * p.centre, nrow = p.centre < 0 ? 9 : 3;
    double g[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) g[c] = 0.0;
    for (int k = lane; k < nc; k += 64) {
      const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
      const double gam =
          2.0 * dl[o.ab + ia * NB + ib] * dl[o.cd + ic * ND + id] -
          0.5 * (dl[o.ac + ia * NC + ic] * dl[o.bd + ib * ND + id] +
                 dl[o.ad + ia * ND + id] * dl[o.bc + ib * NC + ic]);
#pragma unroll
      for (int c = 0; c < 9; ++c)
        if (c >= c0 && c < c0 + nrow) g[c] += gam * output(tb, H, (c - c0) * nc + k);
    }
    // six accumulators per pair (its two atoms' x, y, z), added onto the
    // atoms once every class is done: per-atom atomics would contend
    double* __restrict__ gb = a.grad_pair + 6 * static_cast<size_t>(bidx);
    double* __restrict__ gk = a.grad_pair + 6 * static_cast<size_t>(kidx);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double gA = deg * coop_wave_sum(g[i]);
      const double gB = deg * coop_wave_sum(g[3 + i]);
      const double gC = deg * coop_wave_sum(g[6 + i]);
      if (lane == 0) {
        unsafeAtomicAdd(&gb[i], gA);
        unsafeAtomicAdd(&gb[3 + i], gB);
        unsafeAtomicAdd(&gk[i], gC);
        unsafeAtomicAdd(&gk[3 + i], -(gA + gB + gC));
      }
    }
  }
}

}  // namespace rchem

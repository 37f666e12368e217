// Wave-cooperative contracted ERIs for the classes whose integrals do not fit
// one lane's registers (total angular momentum L >= 5: (dp|pp) ... (dd|dd)),
// for any class on request, and the derivative integrals of the nuclear
// gradient.
//
// One wavefront evaluates one contracted shell quartet at a time, P primitive
// quartets per batch (P a power of two, at most the class's table value, less
// for short contractions).  A workgroup holds W = 1, 2 or 4 such waves, each on
// its own quartets with its own share of LDS; with W > 1 (or on request) the
// VRR operation and accumulation tables are staged in LDS once per workgroup,
// so the operation stream is read from LDS instead of L2, and the waves
// synchronise with wave-level barriers only.  Per wave:
//   1. lanes 0..P-1 set up the batch's primitive-quartet records (exponent
//      factors, PA, WP, QC, WQ, 2 alpha of three centres and
//      G_m = K_ab K_cd (zeta+eta)^(-1/2) F_m(T), m = 0..L) field-major in LDS;
//   2. the Head-Gordon--Pople vertical recurrence runs as the class's
//      operation table on V[slot][q]: lane l always serves primitive quartet
//      q = l mod P and takes operations l / P, l / P + 64 / P, ... of a group,
//      a group being operations of one dependency level and one (side, axis)
//      whose recurrence coefficients the lane loads once; a level boundary is a
//      barrier of the wave;
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
#include "eri_coop.hpp"
#include "types.hpp"

namespace rchem {

namespace {

// Record fields of one primitive quartet (stored field-major: rec[f * P + q])
constexpr int kPA = 0, kWP = 3, kQC = 6, kWQ = 9, kOo2z = 12, kOo2e = 13, kOo2ze = 14,
              kRoz = 15, kRoe = 16, kWA = 17, kWB = 18, kWC = 19, kG = 20;

__host__ __device__ __forceinline__ int rec_fields(int L) { return kG + L + 1; }


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
// H = R; the outputs are read by the epilogue.  vops, vgroups, accs: the
// program's VRR and accumulation tables (device memory or LDS).
template <int BM, int W>
__device__ void quartet(const EriArgs& a, const GenericProgram& p, const uint4* __restrict__ vops,
                        const uint4* __restrict__ vgroups, const uint2* __restrict__ accs,
                        const CoopTables& tbl, const PairRec& bp, const PairRec& kp,
                        double* __restrict__ rec, double* __restrict__ R, int lane) {
  const int L = p.L;
  wsync<W>();  // the previous quartet's outputs have been read
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
  double* __restrict__ Vq = V + q;  // slot s of this lane's primitive quartet: Vq[s << LP]
  const int ostep = 64 >> LP;
  for (int b0 = 0; b0 < npq; b0 += P) {
    wsync<W>();  // the previous batch's records and slots are consumed
    if (lane < P) {
      const int j = b0 + lane;
      const bool valid = j < npq;
      setup_record<BM>(a, bp, kp, valid ? j / nk : 0, valid ? j % nk : 0, valid, L, P, lane, rec);
    }
    wsync<W>();
    for (int idx = lane; idx < (L + 2) * P; idx += 64) {
      const int s = idx >> LP, qq = idx & (P - 1);
      V[idx] = s == 0 ? 0.0 : rec[((kG + s - 1) << LP) + qq];
    }
    wsync<W>();
    for (int g = 0; g < p.nvgroups; ++g) {
      const uint4 gr = vgroups[g];
      const int side = gr.z & 1, ax = (gr.z >> 1) & 3;
      const double c0 = rec[(((side ? kQC : kPA) + ax) << LP) + q];
      const double c1 = rec[(((side ? kWQ : kWP) + ax) << LP) + q];
      const double o = rec[((side ? kOo2e : kOo2z) << LP) + q];
      const double r = rec[((side ? kRoe : kRoz) << LP) + q];
      const double oz = rec[(kOo2ze << LP) + q];
      // the operations of a group are independent (one dependency level):
      // two per iteration, both operation words and all ten operands in
      // flight before either result is formed
      auto vrr_op = [&](const uint4 w) __attribute__((always_inline)) {
        const double n1 = static_cast<double>(w.w & 15u);
        const double n2 = static_cast<double>((w.w >> 4) & 15u);
        double v = c0 * Vq[(w.x >> 16) << LP] + c1 * Vq[(w.y & 0xffffu) << LP];
        v += (n1 * o) * (Vq[(w.y >> 16) << LP] - r * Vq[(w.z & 0xffffu) << LP]);
        v += (n2 * oz) * Vq[(w.z >> 16) << LP];
        return v;
      };
      for (int op = static_cast<int>(gr.x) + (lane >> LP); op < static_cast<int>(gr.y);
           op += 2 * ostep) {
        const int op2 = op + ostep;
        const bool two = op2 < static_cast<int>(gr.y);
        const uint4 w = vops[op];
        const uint4 w2 = vops[two ? op2 : op];
        const double v = vrr_op(w);
        const double v2 = vrr_op(w2);
        Vq[(w.x & 0xffffu) << LP] = v;
        if (two) Vq[(w2.x & 0xffffu) << LP] = v2;
      }
      if (gr.z & 8u) wsync<W>();
    }
    for (int t = lane; t < p.naccs; t += 64) {
      const uint2 e = accs[t];
      const double* src = V + ((e.x & 0xffffu) << LP);
      const int set = static_cast<int>(e.x >> 16);
      double s = 0.0;
      if (set == 0) {
        for (int qq = 0; qq < P; ++qq) s += src[qq];
      } else {
        const double* wt = rec + ((kWA + set - 1) << LP);
        for (int qq = 0; qq < P; ++qq) s += wt[qq] * src[qq];
      }
      acc[e.y] += s;
    }
  }
  wsync<W>();
  coop_hrr<W>(tbl, p.nhgroups, R, bp.AB[0], bp.AB[1], bp.AB[2], kp.AB[0], kp.AB[1], kp.AB[2],
              lane);
}


// Byte size of the VRR and accumulation tables when staged in LDS.
__host__ __device__ inline size_t table_bytes(const GenericProgram& p) {
  return ((16 * static_cast<size_t>(p.nvops) + 16 * static_cast<size_t>(p.nvgroups) +
           8 * static_cast<size_t>(p.naccs)) + 15) & ~static_cast<size_t>(15);
}

// LDS doubles of one wave for program p at batch size P.
__host__ __device__ inline size_t wave_doubles(const GenericProgram& p, int P) {
  const size_t nf = static_cast<size_t>(rec_fields(p.L));
  const size_t buf = std::max(static_cast<size_t>(p.nacc) + static_cast<size_t>(p.nv) * P,
                              static_cast<size_t>(p.nh));
  return static_cast<size_t>(P) * nf + buf + kCoopDBlock;
}

// W waves per workgroup, each on its own quartets; TL: VRR and accumulation
// tables in LDS, and with stage_coop the HRR and output tables too.
template <int MODE, int BM, int W, bool TL>
__global__ __launch_bounds__(64 * W) void eri_generic_kernel(const EriArgs a,
                                                             const GenericProgram p,
                                                             int stage_coop) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint4* vops = p.vops;
  const uint4* vgroups = p.vgroups;
  const uint2* accs = p.accs;
  size_t tb = 0;
  if constexpr (TL) {
    uint4* tv = reinterpret_cast<uint4*>(lds);
    uint4* tg = tv + p.nvops;
    uint2* ta = reinterpret_cast<uint2*>(tg + p.nvgroups);
    for (int i = threadIdx.x; i < p.nvops; i += 64 * W) tv[i] = p.vops[i];
    for (int i = threadIdx.x; i < p.nvgroups; i += 64 * W) tg[i] = p.vgroups[i];
    for (int i = threadIdx.x; i < p.naccs; i += 64 * W) ta[i] = p.accs[i];
    __syncthreads();
    vops = tv;
    vgroups = tg;
    accs = ta;
    tb = table_bytes(p);
  }
  const bool sc = TL && stage_coop != 0;
  const CoopTables tbl = coop_tables(p, reinterpret_cast<char*>(lds) + tb, sc);
  if (sc) {
    tb += coop_table_bytes(p);
    __syncthreads();
  }
  double* rec = reinterpret_cast<double*>(reinterpret_cast<char*>(lds) + tb) +
                static_cast<size_t>(wave) * wave_doubles(p, p.P);
  double* R = rec + p.P * rec_fields(p.L);
  double* dl = rec + wave_doubles(p, p.P) - kCoopDBlock;
  int begin, end;
  if (a.tile_range) {
    begin = a.tile_range[0];
    end = a.tile_range[1];
  } else {
    begin = 0;
    end = a.ntiles;
  }
  // every wave leaves once its next tile index passes the end; the waves of a
  // workgroup share nothing after the table load, so they may leave apart
  for (int ti = begin + static_cast<int>(blockIdx.x) * W + wave; ti < end;
       ti += static_cast<int>(gridDim.x) * W) {
    const Tile t = a.tiles[ti];
    for (int i = 0; i < t.count; ++i) {
      const int kidx = t.ket0 + i;
      const int bidx = (MODE == kModeDiag) ? kidx : t.bra;
      const PairRec bp = a.pairs[bidx];
      const PairRec kp = a.pairs[kidx];
      quartet<BM, W>(a, p, vops, vgroups, accs, tbl, bp, kp, rec, R, lane);
      epilogue<W, MODE>(a, p, tbl, t, bidx, kidx, bp, kp, R, dl, lane);
    }
  }
}

template <int MODE, int BM, int W, bool TL>
hipError_t launch_shape(const EriArgs& a, const GenericProgram& p, int sc, unsigned grid,
                        size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((eri_generic_kernel<MODE, BM, W, TL>), dim3(grid), dim3(64 * W), lds, s, a,
                     p, sc);
  return hipGetLastError();
}

template <int MODE, int BM>
hipError_t launch_mode(const EriArgs& a, const GenericProgram& p, int W, bool tl, int sc,
                       unsigned grid, size_t lds, hipStream_t s) {
  if (W == 4) return launch_shape<MODE, BM, 4, true>(a, p, sc, grid, lds, s);
  if (W == 2) return launch_shape<MODE, BM, 2, true>(a, p, sc, grid, lds, s);
  return tl ? launch_shape<MODE, BM, 1, true>(a, p, sc, grid, lds, s)
            : launch_shape<MODE, BM, 1, false>(a, p, 0, grid, lds, s);
}

}  // namespace

// LDS bytes of one one-wave workgroup for program p with its tables in
// device memory (the smallest footprint: whether the class fits a workgroup).
size_t eri_generic_lds_bytes(const GenericProgram& p) {
  return sizeof(double) * wave_doubles(p, p.P);
}

// Launch shapes (the tuning table's variant for a class on this kernel):
// {waves per workgroup, tables in LDS, batch divisor}; shape 0 picks the
// largest W whose workgroup fits the device's LDS with the tables staged.
struct GenericShape {
  int waves, tables_in_lds, pdiv;
};
constexpr GenericShape kGenericShapes[] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 1}, {2, 1, 1},
                                           {4, 1, 1}, {2, 1, 2}, {4, 1, 2}, {4, 1, 4}};
constexpr int kNumGenericShapes = sizeof(kGenericShapes) / sizeof(kGenericShapes[0]);

int eri_generic_shapes() { return kNumGenericShapes; }

// blocks: with a device tile range, the number of 4-wave blocks the per-lane
// kernels would use (this kernel launches as many waves, W per workgroup);
// without one, one wave per tile.
hipError_t eri_generic_launch(const EriArgs& a, const GenericProgram& prog, int mode, int bm,
                              int blocks, hipStream_t s, size_t lds_limit, int shape) {
  GenericProgram p = prog;
  if (p.P < 1 || p.P > 64 || (p.P & (p.P - 1)) != 0 || p.L > kBoysMaxM) return hipErrorInvalidValue;
  if (shape < 0 || shape >= kNumGenericShapes) shape = 0;
  GenericShape sh = kGenericShapes[shape];
  auto bytes = [&](int W, bool tl, int P) {
    return (tl ? table_bytes(p) : 0) + sizeof(double) * W * wave_doubles(p, P);
  };
  if (sh.waves == 0) {
    sh = {1, 0, 1};
    for (int W : {4, 2, 1})
      if (bytes(W, true, p.P) <= lds_limit) {
        sh = {W, 1, 1};
        break;
      }
  }
  int P = std::max(1, p.P / sh.pdiv);
  if (bytes(sh.waves, sh.tables_in_lds != 0, P) > lds_limit) {  // shape too big: one wave
    sh = {1, 0, 1};
    P = p.P;
  }
  p.P = P;
  size_t lds = bytes(sh.waves, sh.tables_in_lds != 0, P);
  if (lds > lds_limit) return hipErrorInvalidValue;
  // the HRR and output tables too, where they fit beside the rest
  int sc = 0;
  if (sh.tables_in_lds && lds + coop_table_bytes(p) <= lds_limit) {
    sc = 1;
    lds += coop_table_bytes(p);
  }
  unsigned waves;
  if (a.tile_range) {
    if (blocks <= 0) return hipSuccess;
    waves = static_cast<unsigned>(blocks) * 4u;
  } else {
    if (a.ntiles <= 0) return hipSuccess;
    waves = static_cast<unsigned>(a.ntiles);
  }
  const unsigned grid = (waves + sh.waves - 1) / sh.waves;
  const bool tl = sh.tables_in_lds != 0;
  const int W = sh.waves;
  const bool compat = bm == kBoysCompat;
  switch (mode) {
    case kModeJK:
      return compat ? launch_mode<kModeJK, kBoysCompat>(a, p, W, tl, sc, grid, lds, s)
                    : launch_mode<kModeJK, kBoysExact>(a, p, W, tl, sc, grid, lds, s);
    case kModeDiag:
      return compat ? launch_mode<kModeDiag, kBoysCompat>(a, p, W, tl, sc, grid, lds, s)
                    : launch_mode<kModeDiag, kBoysExact>(a, p, W, tl, sc, grid, lds, s);
    case kModeTensor:
      return compat ? launch_mode<kModeTensor, kBoysCompat>(a, p, W, tl, sc, grid, lds, s)
                    : launch_mode<kModeTensor, kBoysExact>(a, p, W, tl, sc, grid, lds, s);
    case kModeQuartet:
      return compat ? launch_mode<kModeQuartet, kBoysCompat>(a, p, W, tl, sc, grid, lds, s)
                    : launch_mode<kModeQuartet, kBoysExact>(a, p, W, tl, sc, grid, lds, s);
    case kModeGrad:
      return launch_mode<kModeGrad, kBoysExact>(a, p, W, tl, sc, grid, lds, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace rchem

// Data layout shared by the host library and the gfx950 kernels.
//
// Shell pairs are the unit of work on both sides of a quartet.  They live in
// one array ordered by (pair class, primitive-pair count, Schwarz bound
// descending).  A run of pairs with the same class and the same primitive-pair
// count is a "bucket"; its primitive-pair records are stored transposed
// ([primitive pair][pair in bucket]) so that the 64 lanes of a wavefront, which
// hold 64 consecutive ket pairs of one bucket, read consecutive records.
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace rchem {

// Primitives per shell accepted by the library (a pair has at most 64
// primitive pairs).
constexpr int kMaxPrimPerShell = 8;
constexpr int kMaxPrimPairs = kMaxPrimPerShell * kMaxPrimPerShell;

// One significant shell pair (a, b); a is the shell with the higher l.
struct PairRec {
  double A[3];   // centre of shell a
  double AB[3];  // A - B
  int fa, fb;    // first basis-function index of shells a and b
  int np;        // number of primitive pairs
  int p0;        // index of primitive pair 0 in the primitive-pair array
  int ps;        // stride between consecutive primitive pairs of this pair
  int flags;     // bit 0: a and b are different shells
};

// One primitive pair (i of a, j of b).  The product centre is stored relative
// to shell a's centre, P - A = -(alpha_j / zeta) (A - B), so that no lane
// forms it by cancellation, and P - Q = (P - A) - (Q - C) + (A - C).
struct PrimRec {
  double zeta;          // alpha_i + alpha_j
  double pax, pay, paz; // P - A
  double kab;           // c_i N_i c_j N_j sqrt(2) pi^(5/4) / zeta exp(-alpha_i alpha_j/zeta |AB|^2)
  double oo2z;          // 1 / (2 zeta)
};

// The same record rounded to FP32 (same index), for the FP32 kernels: half
// the bytes per primitive pair.
struct PrimRecF {
  float zeta, pax, pay, paz, kab, oo2z;
};

// A wavefront's work: one bra pair against up to 64 consecutive ket pairs of
// one bucket (uniform primitive-pair count across the wave).
struct Tile {
  int bra;
  int ket0;
  int count;
  int pad;
};

// A block's work in LDS-digestion mode: tiles [tb, te) of one bra pair.
struct Chunk {
  int bra;
  int tb, te;
  int pad;
};

enum EriMode { kModeJK = 0, kModeDiag = 1, kModeTensor = 2, kModeQuartet = 3, kModeGrad = 4 };

// Launch shapes of the LDS-digestion kernel a class can be compiled in (the
// per-class tuning table picks one): waves per workgroup (they share the bra's
// K rows), the minimum waves per SIMD the register allocation must allow (0:
// no bound), whether a lane fetches its next ket primitive pair one iteration
// ahead (latency hiding in registers instead of by other waves), how many bra
// primitive pairs one iteration of the inner loop takes (2: two independent
// dependency chains to interleave), and whether the Boys function of orders
// 0 and 1 is evaluated without a branch (both the table and the asymptotic
// form, then a select: one basic block, so the unrolled chains interleave).
struct LdsVariant {
  int waves, min_waves_per_simd, prefetch, unroll, branch_free;
};
constexpr LdsVariant kLdsVariants[] = {
    {4, 0, 1, 1, 0}, {4, 0, 0, 1, 0}, {8, 0, 1, 1, 0}, {8, 0, 0, 1, 0}, {8, 6, 0, 1, 0},
    {8, 8, 0, 1, 0}, {16, 0, 0, 1, 0}, {8, 0, 0, 2, 1}, {4, 0, 0, 2, 1}, {8, 0, 0, 1, 1}};
constexpr int kNumLdsVariants = sizeof(kLdsVariants) / sizeof(kLdsVariants[0]);

// Pair-indexed Coulomb accumulators: component k of pair p (pair class X) is
// Jp[base_X + k * n_X + (p - begin_X)], so the 64 lanes of a wave, which hold
// consecutive ket pairs, update consecutive addresses.
struct JpClass {
  int base, n, begin, pad;
};

struct EriArgs {
  const PairRec* pairs;
  const PrimRec* prims;
  const PrimRecF* prims_f;  // prims rounded to FP32
  const Tile* tiles;
  int ntiles;              // host-sized launch: tiles[0 .. ntiles)
  const int* tile_range;   // if set (persistent launch): tiles[range[0] .. range[1]),
                           // read from device memory, drained by a fixed grid
  int nbf;
  const double* boys_table;
  const float* boys_table_f;
  const double* D;   // kModeJK: density, nbf x nbf row-major, symmetric
  double* J;         // kModeJK: lower-half accumulators (J + J^T is the result)
  double* K;
  double* out;       // kModeDiag: Schwarz bound per pair; kModeTensor: nbf^4 tensor;
                     // kModeQuartet: NOUT values per tile at out[tile.pad * NOUT]
  // LDS digestion (chunked launches): K rows of the bra in LDS, J by pair
  const Chunk* chunks;
  const int* chunk_range;  // [begin, end) in chunks, device memory
  double* Jp;
  JpClass jp_bra, jp_ket;
  const double* Dp;  // the density blocks of the pairs, in Jp's layout (null: gather from D)
  // K rows in LDS span columns [kc0, kc0 + kw) of the bra's rows: kc0 = 0 and
  // kw = nbf for one molecule; in a multi-molecule basis (quartets within a
  // molecule) kc0 = pair_kc0[bra], the first function of the bra's molecule,
  // and kw the largest molecule's function count
  int kw;
  const int* pair_kc0;
  // kModeGrad (wave-cooperative kernel): exponent of the first shell of each
  // primitive pair (same layout as prims), atoms of each pair's two shells,
  // and the 3 x natom gradient accumulated into
  const double* prim_alpha;
  const int2* pair_atoms;
  double* grad;
  // per-lane gradient kernel: 6 accumulators per pair (its two atoms x, y, z),
  // reduced onto the atoms afterwards (spreads the atomics over many addresses)
  double* grad_pair;
};

// A class served by the wave-cooperative kernel: its recurrences as operation
// tables in device memory (layouts in gen_eri_prog.py).  One wavefront
// evaluates one contracted quartet; its share of LDS holds the records of a
// batch of P primitive quartets and one buffer of max(nacc + nv P, nh) doubles
// that is the contracted accumulators + VRR slots during the primitive loop
// and the HRR workspace afterwards.  A workgroup of several waves (each on its
// own quartets) can stage the VRR and accumulation tables in LDS once.
struct GenericProgram {
  int la, lb, lc, ld;
  int L;          // highest Boys order (la + lb + lc + ld, + 1 for the gradient)
  int nacc;       // contracted accumulators
  int nv;         // VRR slots per primitive quartet (slot 0 = zero, 1..L+1 = G_m)
  int nh;         // HRR slots (the first nacc are the accumulators)
  int nout;       // outputs (integrals, or 9 x integrals for the gradient)
  int nvops, nvgroups, nhops, nhgroups, naccs;
  int P;          // primitive quartets per batch (power of two)
  int centre;     // gradient program of one centre (0 A, 1 B, 2 C; outputs 3 x components),
                  // -1: all three (outputs 9 x components) or not a gradient program
  const uint4* vops;      // VRR operations
  const uint4* vgroups;   // (op_begin, op_end, side | axis << 1 | level_end << 3, 0)
  const uint2* accs;      // (slot | set << 16, accumulator)
  const uint2* hops;      // HRR operations
  const uint4* hgroups;
  const int4* outs;       // (plus slot, minus slot, n, 0): norm (H[plus] - n H[minus])
  const double* out_norm;
};

// Launcher of a class's streaming kernel (eri_stream.hpp) in one mode: the
// class's wave-cooperative program supplies the HRR and output tables.
using EriStreamFn = hipError_t (*)(const EriArgs&, const GenericProgram&, int blocks,
                                   hipStream_t, size_t lds_limit);

// The planner's per-class statistics are summed into this many copies (a
// wave adds into copy wave % kPlanStatsReplicas), so that the adds of the
// ~10^4 waves of a build do not all queue on the same few addresses; readers
// sum the copies.
constexpr int kPlanStatsReplicas = 64;

// One quartet class as the device planner sees it.
struct PlanClass {
  int bra_begin, bra_end;  // bra pairs (pair class X)
  int bk_off, nbk;         // ket buckets (pair class Y) in the bucket array
  int task_begin;          // first task; tasks = (bra - bra_begin) * nbk + bucket
  int same;                // X == Y: keep only ket <= bra
  int unit_begin;          // first bra unit; units = bra - bra_begin
  int chunk64, chunk32;    // tiles per LDS-digestion chunk, FP64 and FP32 lists
  int order;               // bit 0 / 1: FP64 / FP32 chunks ordered by first ket pair
};

struct PlanArgs {
  const PlanClass* classes;
  int nclasses;
  int ntasks;
  const int4* buckets;     // {begin, end, np, 0}
  const double* Q;         // Schwarz bound per pair, layout order
  const PairRec* pairs;
  double tau;              // Schwarz threshold
  double mixed_thresh;     // FP32 below Q_ab Q_cd max|D| < mixed_thresh; <= 0: all FP64
  const double* dmax;      // max |D| (device scalar)
  int* cnt64;              // per task: tiles, then (after the scan) offsets; length ntasks + 1
  int* cnt32;
  int* lo;                 // per task ket range [lo, mid) in FP64, [mid, hi) in FP32
  int* mid;
  int* hi;
  double* stats;           // per class: {quartets64, quartets32, prim64, prim32}, in
                           // kPlanStatsReplicas replicas stats_stride doubles apart
  int stats_stride;
  int nunits;              // bra units (class, bra pair)
  // several molecules in one basis: kets restricted to the bra's molecule,
  // segs[(bucket slot) * nmol + molecule] = its pairs [begin, end) in the bucket
  const int2* segs;        // null: one molecule
  int nmol;
  const int* pair_mol;     // molecule of each pair (layout order)
  const int* mol_active;   // per molecule: 0 retired (no quartets), 1 active, 2 active with
                           // every quartet in FP64 (past its precision switch); null: all 1
};

// Device buffers the planner fills (all sized by the host from upper bounds).
struct PlanBuffers {
  int* off64;        // tile offsets per task (exclusive scan), ntasks + 1
  int* off32;
  Tile* tiles64;     // cap entries each
  Tile* tiles32;
  int* chc64;        // chunk counts per unit, nunits + 1
  int* chc32;
  int* choff64;      // chunk offsets per unit (exclusive scan)
  int* choff32;
  Chunk* chunks64;   // cap entries each
  Chunk* chunks32;
  int cap;
  int* ranges;       // 8 per class
  void* scan_tmp;
  size_t scan_bytes;
  // chunk order by ket window (key_bits > 0): the chunks are written unsorted
  // with a key (class, first ket pair / 64) and radix-sorted into chunks64 /
  // chunks32, so that the workgroups running at one time read the same ket
  // pairs (their records stay in L2) rather than every bra's whole ket range
  int key_bits;      // 0: bra order (no sort)
  int sort_mask;     // bit 0 / 1: sort the FP64 / FP32 chunk list
  int ccap;          // chunks sorted per precision (upper bound of the chunk count)
  Chunk* chunks_raw;
  unsigned* keys_raw;
  unsigned* keys_sorted;
  void* sort_tmp;
  size_t sort_bytes;
};

// Device-side state of one DIIS SCF (Pulay extrapolation over a ring of at most
// kDiisMax Fock matrices and their orthogonalised commutators), updated by the
// SCF step's kernels so that a whole iteration can be replayed as one graph.
// The leading scalars are what the host reads back after an iteration.
constexpr int kDiisMax = 32;
struct ScfDevState {
  double e_cur;       // sum D o (H + F(D)) of the density this iteration's F was built from
  double de;          // e_cur - e_prev
  double errmax;      // max |X^T (FDS - SDF) X|
  int converged;      // |de| < thresh_e, errmax < errtol, all-FP64 build, not the first iteration
  int switched;       // mixed build switches to all-FP64: D kept, nothing extrapolated
  int singular;       // DIIS system singular: history restarted, next D from F
  int purify_failed;  // fixed-step purification ended above its tolerance: D kept
  int it;             // iterations done
  int count;          // commutators pushed since the last restart
  int m;              // subspace size of this iteration
  int slot;           // ring slot of this iteration's push
  int damp;           // this iteration's density is damped
  int purify_steps;   // steps the fixed-step purification needed (first below tolerance + 3)
  double e_prev;
  double c[kDiisMax];               // extrapolation coefficients (ring order)
  double B[kDiisMax * kDiisMax];    // commutator dot products (ring order)
};
struct ScfDevParams {
  double thresh_e, errtol;     // convergence: |dE| and max |commutator|
  double switch_de;            // mixed -> all-FP64 once |dE| falls below this
  double damping, damping_until;
  int k;                       // ring length (<= kDiisMax)
  int mixed;                   // the build being replayed is the mixed one
  int fp64;                    // the build being replayed is the all-FP64 one
  int pad;
};
constexpr int kDiisBlocks = 240;  // fixed grid of the dot-product pass: fixed summation order

}  // namespace rchem

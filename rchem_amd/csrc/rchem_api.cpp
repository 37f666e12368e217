// Host side of the C ABI declared in include/rchem_gpu.h.
//
// Owns the context (device, stream, rocBLAS/rocSOLVER handles), the basis in
// shell form, the shell-pair and tile lists that drive the ERI kernels, and
// the SCF driver.  All arithmetic on integrals and matrices runs on the GPU;
// the host sorts pair lists and drives launches.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/rchem_gpu.h"
#include "boys.hpp"
#include "types.hpp"

namespace rchem {
#include "generated/basis_tables.inc"
#include "generated/eri_flops.inc"

using EriLaunchFn = hipError_t (*)(const EriArgs&, int, hipStream_t);
EriLaunchFn eri_launcher(int la, int lb, int lc, int ld, int mode, int bm, int fp32);
hipError_t dmax_launch(const double* D, size_t n2, unsigned long long* out, hipStream_t s);
hipError_t dmax_blocks_launch(const double* D, size_t n, const int* f0, const int* nb, int nmol,
                              int nmax, unsigned long long* out, hipStream_t s);
size_t plan_scan_bytes(int ntasks);
size_t plan_sort_bytes(int n);
int plan_class_bits(int nclasses);
hipError_t plan_launch(const PlanArgs& a, const PlanBuffers& b, hipStream_t s);
EriLaunchFn eri_jk_lds_launcher(int la, int lb, int lc, int ld, int bm, int fp32, int v);
EriStreamFn eri_stream_launcher(int la, int lb, int lc, int ld, int mode);
hipError_t eri_generic_launch(const EriArgs& a, const GenericProgram& p, int mode, int bm,
                              int blocks, hipStream_t s, size_t lds_limit, int shape);
int eri_generic_shapes();

size_t eri_generic_lds_bytes(const GenericProgram& p);

// Operation tables of the wave-cooperative kernel (host copies): the integrals
// of every class, and the derivative integrals of the gradient (grad = 1, or
// one program per centre, grad = 2 + centre, where the full one is too large).
struct HostProgram {
  int la, lb, lc, ld, grad, L, nacc, nv, nh, nout, nvops, nvgroups, naccs, nhops, nhgroups, P;
  const uint32_t* vops;
  const uint32_t* vgroups;
  const uint32_t* accs;
  const uint32_t* hops;
  const uint32_t* hgroups;
  const int* outs;
  const double* out_norm;
};
#include "generated/eri_programs.inc"
constexpr int kNumHostPrograms = sizeof(kHostPrograms) / sizeof(kHostPrograms[0]);
hipError_t dpair_gather_launch(const PairRec* pairs, const double* D, const JpClass* cls,
                               const int* ncomp, int ncls, int npairs, double* Dp, int nbf,
                               hipStream_t s);
hipError_t jp_scatter_launch(const PairRec* pairs, const double* Jp, const JpClass* cls,
                             const int* ncomp, int ncls, int npairs, double* Jh, int nbf,
                             hipStream_t s);
hipError_t symmetrize_launch(const double* Jh, const double* Kh, double* J, double* K, int n,
                             hipStream_t s);
hipError_t symmetrize_blocks_launch(const double* Jh, const double* Kh, double* J, double* K,
                                    int n, const int* f0, const int* nb, int nmol, int nmax,
                                    hipStream_t s);
hipError_t fill_blocks_launch(double* A, double* B, int n, const int* f0, const int* nb, int nmol,
                              int nmax, hipStream_t s);
hipError_t fill_launch(double* p, size_t n, double v, hipStream_t s);
hipError_t fock_launch(const double* H, const double* J, const double* K, double* F, size_t n2,
                       hipStream_t s);
hipError_t energy_launch(const double* D, const double* H, const double* F, double* out,
                         size_t n2, hipStream_t s);
hipError_t antisym_launch(const double* A, double* E, int n, hipStream_t s);
hipError_t purify_start_launch(const double* F, int n, int nocc, double* lo, double* hi,
                               double* par, double* P, hipStream_t s);
hipError_t purify_step_launch(double* P, const double* P2, const double* P3, int n,
                              double* coef, hipStream_t s);
hipError_t diis_push_launch(const double* E, const double* F, double* Es, double* Fs, size_t n2,
                            const ScfDevState* st, int k, double* part, hipStream_t s);
hipError_t diis_solve_launch(ScfDevState* st, const double* part, const double* energy,
                             const ScfDevParams& p, hipStream_t s);
hipError_t diis_extrapolate_launch(const double* Fs, size_t n2, const ScfDevState* st, double* Fx,
                                   hipStream_t s);
hipError_t scf_commit_launch(double* D, const double* Dn, size_t n2, ScfDevState* st,
                             const double* coef, int steps, double tol, double damping,
                             hipStream_t s);
hipError_t batch_pack(const double* A, size_t n, const int* f0, const int* nb, int nmax, int nbatch,
                      double pad, double* out, hipStream_t s);
hipError_t batch_unpack(const double* in, int nmax, const int* f0, const int* nb, int nbatch,
                        size_t n, double* A, hipStream_t s);
hipError_t batch_mask_cols(double* C, int nmax, const int* nocc, int nbatch, hipStream_t s);
hipError_t batch_inv_sqrt_cols(const double* U, const double* w, int nmax, int nbatch, double* out,
                               hipStream_t s);
hipError_t batch_energy(const double* D, const double* H, const double* F, size_t n, const int* f0,
                        const int* nb, int nbatch, double* out, hipStream_t s);
hipError_t batch_antisym(const double* A, double* E, int nmax, int nbatch, hipStream_t s);
hipError_t batch_dots(const double* base, const long long* pairs, int npairs, size_t len,
                      double* out, hipStream_t s);
hipError_t batch_maxabs(const double* E, int nmax, int nbatch, double* out, hipStream_t s);
hipError_t batch_blend(double* D, const double* Dold, size_t n, const int* f0, const int* nb,
                       const double* alpha, int nmax, int nbatch, hipStream_t s);
hipError_t batch_combine(const double* Fs, size_t batch_stride, int m, int k, const double* c,
                         int nmax, int nbatch, double* out, hipStream_t s);
hipError_t batch_gather(const double* in, size_t in_stride, int nin, const int* idx, int nout_batch,
                        int nring, double* out, size_t out_stride, int nout, hipStream_t s);
hipError_t batch_fock(const double* H, const double* J, const double* K, double* F, size_t n,
                      const int* f0, const int* nb, int nmol, int nmax, hipStream_t s);
hipError_t scale_columns_launch(const double* U, const double* w, double* out, int n,
                                hipStream_t s);

struct ShellDev {
  double xyz[3];
  int l, nprim, p0, f0;
};
struct OneElArgs {
  const ShellDev* shells;
  const double* exps;
  const double* ccoef;
  const int2* pairs;
  int npairs;
  int nbf;
  int natom;
  const double* atom_xyz;
  const double* atom_z;
  const double* boys_table;
  double* S;
  double* T;
  double* V;
  const int* shell_mol;   // several molecules: the molecule of each shell, else null
  const int2* mol_atoms;  // nuclei [begin, end) of each molecule
  int cls_begin[10];      // pairs sorted by class 3 la + lb: class c is [cls_begin[c], cls_begin[c + 1])
};
// Sorts shell pairs by class 3 l_a + l_b (stable) and records the class
// boundaries, the order one_electron_launch expects.
void sort_pairs_by_class(const std::vector<ShellDev>& sh, std::vector<int2>* pr, int* cls_begin) {
  auto cls = [&](const int2& q) { return 3 * sh[q.x].l + sh[q.y].l; };
  std::stable_sort(pr->begin(), pr->end(),
                   [&](const int2& x, const int2& y) { return cls(x) < cls(y); });
  size_t i = 0;
  for (int c = 0; c < 10; ++c) {
    while (i < pr->size() && cls((*pr)[i]) < c) ++i;
    cls_begin[c] = static_cast<int>(i);
  }
}
struct OneElGradArgs {
  const ShellDev* shells;
  const double* exps;
  const double* ccoef;
  const int2* pairs;
  int npairs;
  int nbf;
  int natom;
  const double* atom_xyz;
  const double* atom_z;
  const double* boys_table;
  const int* shell_atom;
  const double* D;
  const double* W;
  double* grad;
  int cls_begin[10];
};
hipError_t one_electron_grad_launch(const OneElGradArgs& a, hipStream_t s);
hipError_t one_electron_launch(const OneElArgs& a, hipStream_t s);
using EriGradFn = hipError_t (*)(const EriArgs&, hipStream_t);
EriGradFn eri_grad_launcher(int la, int lb, int lc, int ld);
hipError_t grad_pairs_launch(const double* gp, const int2* atoms, int npairs, double* grad,
                             hipStream_t s);
}  // namespace rchem

using namespace rchem;

namespace {

constexpr int kMaxL = 2;        // highest angular momentum accepted
constexpr int kNumPairClasses = (kMaxL + 1) * (kMaxL + 2) / 2;

int pair_class(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }
int ncart(int l) { return (l + 1) * (l + 2) / 2; }

double fact2(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, b ? b : 8);
    if (e == hipSuccess) bytes = b ? b : 8;
    return e;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

struct Shell {
  int l = 0, nprim = 0, f0 = 0, p0 = 0;
  double xyz[3] = {0, 0, 0};
};

struct Bucket {
  int begin, end, np;  // pairs [begin, end) of the sorted pair array, np primitive pairs each
};

struct QClass {
  int X, Y;           // bra / ket pair classes
  int la, lb, lc, ld;
  int tile_begin, ntiles;  // host-tiled launches
  int cap;                 // planner: upper bound on tiles of one precision
  double quartets, prim_quartets, flops;  // all-FP64 plan
};

struct ClassTimer {
  int la, lb, lc, ld, fp32;
  int cls;  // index into the planner's classes
  double quartets, prim_quartets, flops;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  int launches = 0;
};

// Launch choices of one (la lb|lc ld) class in one precision of the Fock
// build: LDS-digestion kernel shape (kLdsVariants), persistent workgroups, and
// tiles per chunk.  Defaults: shape 0, 2048 workgroups (256 CUs x 8), 64 tiles.
constexpr int kDefaultBlocks = 2048;
constexpr int kDefaultChunkTiles = 64;
struct ClassTune {
  int la = 0, lb = 0, lc = 0, ld = 0, fp32 = 0;
  int variant = 0;
  int blocks = kDefaultBlocks;
  int chunk_tiles = kDefaultChunkTiles;
  int ket_order = -1;  // LDS chunks in ket order: 1 yes, 0 no, -1 the default (FP64 yes, FP32 no)
};

// Host-side record of a shell pair before it is laid out.
struct PairHost {
  int a, b;     // shell indices, l(a) >= l(b)
  int cls, np;
  double Q;
};

}  // namespace

struct rchem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::string err;

  // basis
  std::vector<Shell> shells;
  std::vector<double> exps, coefs, ccoef;  // raw coefficients and c * radial norm
  int nbf = 0;
  bool have_basis = false;

  // Several independent molecules in one basis (batched SCF): molecule of
  // every shell; only pairs and quartets within one molecule exist.
  int nmol = 1;
  // a multi-molecule basis: each molecule's first function and function
  // count (device copies), the largest count; J and K are block-diagonal
  DevBuf mol_f0_d, mol_nb_d;
  int mol_nmax = 0;
  std::vector<int> shell_mol;                   // empty: one molecule
  std::vector<int> pair_mol;                    // layout order
  std::vector<std::vector<int2>> bucket_segs;   // [pair class][bucket * nmol + molecule]
  DevBuf pair_mol_d, plan_segs_d;
  // batched SCF: per-molecule flags the planner reads (0 = converged and
  // retired, its quartets skipped); used while mol_active_on
  DevBuf dp_d;  // pair-indexed density blocks (layout of jp_d), rebuilt every Fock build
  DevBuf mol_active_d;
  bool mol_active_on = false;
  // multi-molecule basis: first function of each pair's molecule (layout
  // order) and the largest molecule's function count, the K-row window of
  // LDS digestion
  DevBuf pair_kc0_d;
  int mol_kw = 0;

  // Schwarz bounds of every shell pair (a >= b), computed once per basis
  std::vector<PairHost> all_pairs;
  bool have_schwarz = false;

  // laid-out pair list for a screening threshold
  double pairs_tau = -1.0;
  std::vector<PairRec> pairs;
  std::vector<double> pairQ;                      // Schwarz bound, layout order
  std::vector<std::pair<int, int>> pair_shells;   // (a, b), layout order
  size_t nprims = 0;
  std::vector<Bucket> buckets[kNumPairClasses];
  int class_begin[kNumPairClasses + 1] = {0};
  DevBuf pairs_d, prims_d, prims_f_d;
  DevBuf alpha_d;  // exponent of the first shell of every primitive pair (prims layout)

  // tiles for a threshold
  double tiles_tau = -1.0;
  std::vector<Tile> tiles;
  std::vector<QClass> qclasses;   // device-planner classes
  std::vector<QClass> hclasses;   // classes of the host-built tile list (tensor, gradient)
  DevBuf tiles_d;
  rchem_jk_stats planned{};  // statistics of the planned tile list

  // device planner
  double plan_tau = -1.0;
  // The plan is cut into passes of at most kPassTiles tiles: a pass is a list
  // of segments (a class restricted to a window of bra pairs), planned and
  // launched together; every build runs the passes in order into the same
  // J, K and Coulomb accumulators.  pclasses holds the segments of all passes.
  struct Pass {
    int seg_begin, nseg, tasks, units, cap;
    int ccap;   // chunks per precision, upper bound (tiles / chunk size + units)
    int order;  // bit 0 / 1: some class's FP64 / FP32 chunks in ket order
  };
  std::vector<PlanClass> pclasses;
  std::vector<int> seg_class;  // segment -> qclasses index
  std::vector<int> seg_cap;    // tiles reserved for the segment
  std::vector<Pass> passes;
  int plan_tasks = 0, plan_cap = 0;
  size_t plan_scan_bytes = 0;
  int plan_units = 0;
  DevBuf plan_classes_d, plan_buckets_d, plan_q_d, plan_cnt64, plan_cnt32, plan_off64,
      plan_off32, plan_lo, plan_mid, plan_hi, plan_tiles64, plan_tiles32, plan_ranges,
      plan_stats, plan_dmax, plan_scan_tmp, plan_chc64, plan_chc32, plan_choff64, plan_choff32,
      plan_chunks64, plan_chunks32;
  // LDS chunks in ket order (per class and precision from the tuning;
  // RCHEM_CHUNK_ORDER=0 / 1 turns it off / on for every class): sort keys and buffers
  int chunk_order_env = -1;
  int plan_key_bits = 0, plan_ccap = 0;
  size_t plan_sort_bytes = 0;
  DevBuf plan_chunks_raw, plan_keys_raw, plan_keys_sorted, plan_sort_tmp;
  // pair-indexed Coulomb accumulators
  std::vector<JpClass> jp_cls;
  size_t jp_size = 0;
  DevBuf jp_d, jp_cls_d, jp_ncomp_d;

  DevBuf boys_d, boys_f_d;
  // state of the last SCF, kept for the gradient: the density and the Fock
  // matrix it is the density of (W = D F D)
  DevBuf scf_D, scf_F, scf_Dold;
  ScfDevState* scf_status_h = nullptr;  // pinned: the scalars read after each iteration
  DevBuf work_Jh, work_Kh, work_D, work_J, work_K;
  rchem_jk_stats last_stats{};
  const double* last_stats_dev = nullptr;  // per-class planner statistics of the last build
  bool last_mixed = false;
  bool last_mixed_pending = false;  // the build being launched is mixed

  // LDS per workgroup of the device (hipDeviceProp_t::sharedMemPerBlock)
  size_t lds_block = 65536;
  // per (class, precision) launch choices of the Fock build, from the tuning
  // table (defaults where it has no entry)
  std::vector<ClassTune> tuning;
  std::vector<ClassTune> tuning_batch;  // used in place of tuning by rchem_batch_scf (if any)

  bool profiling = false;
  bool lds_digestion = true;  // RCHEM_LDS_DIGESTION=0 selects global-atomic digestion
  bool force_generic = false;  // RCHEM_ERI_GENERIC=1: every class on the wave-cooperative kernel
  int eri_stream_mode = 1;     // streaming kernel: 0 never, 1 where faster, 2 wherever it serves
  bool progress = false;       // RCHEM_PROGRESS=1: report each finished pass of a Fock build
  bool capturing = false;      // a graph capture is recording the stream (no synchronisation)
  // wave-cooperative kernel: operation tables in device memory (uploaded on first use)
  std::vector<GenericProgram> gprog;
  DevBuf gprog_d;
  std::vector<ClassTimer> timers;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> event_pool;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
};

static std::string g_create_error;

namespace rchem {
// error text reported by rchem_last_error(NULL) (context creation and the
// context-free compatibility functions)
void set_global_error(const std::string& msg) { g_create_error = msg; }
}  // namespace rchem

#define HIPC(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ctx->fail(RCHEM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define BLASC(expr)                                                                       \
  do {                                                                                    \
    rocblas_status s_ = (expr);                                                           \
    if (s_ != rocblas_status_success)                                                     \
      return ctx->fail(RCHEM_ERR_HIP, std::string(#expr) + ": rocblas status " +          \
                                          std::to_string(static_cast<int>(s_)));          \
  } while (0)

#define TRY(expr)                \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != RCHEM_OK) return rc_; \
  } while (0)

namespace {

// Boys tables: for each order M, row r holds F_{M+j}(r/2) / j!,
// j = 0..kBoysTerms-1, from the all-positive series
// F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) in long double;
// then exp(-r/2) for every row.
std::vector<double> make_boys_table() {
  constexpr int kM = kBoysMaxM + kBoysTerms;  // orders needed
  std::vector<double> tab(static_cast<size_t>(kBoysTableSize));
  std::vector<long double> F(kM);
  for (int r = 0; r < kBoysRows; ++r) {
    const long double T = static_cast<long double>(r) / kBoysInvH;
    for (int m = 0; m < kM; ++m) {
      long double term = 1.0L / (2 * m + 1), sum = term;
      for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * T / (2 * m + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
      }
      F[m] = expl(-T) * sum;
    }
    for (int M = 0; M <= kBoysMaxM; ++M) {
      long double fact = 1.0L;
      for (int j = 0; j < kBoysTerms; ++j) {
        if (j > 0) fact *= j;
        tab[static_cast<size_t>(M) * kBoysBlock + r * kBoysTerms + j] =
            static_cast<double>(F[M + j] / fact);
      }
    }
    tab[kBoysExpOffset + r] = static_cast<double>(expl(-T));
  }
  return tab;
}

// The tuning entry of class (la lb|lc ld) in one precision, or the defaults.
ClassTune tune_for(const rchem_ctx* ctx, int la, int lb, int lc, int ld, int fp32) {
  for (const ClassTune& t : ctx->tuning)
    if (t.la == la && t.lb == lb && t.lc == lc && t.ld == ld && t.fp32 == fp32) return t;
  ClassTune d;
  d.la = la; d.lb = lb; d.lc = lc; d.ld = ld; d.fp32 = fp32;
  return d;
}

// Whether the class's LDS chunks of one precision are planned in ket order:
// concurrently running workgroups then read the same ket pairs (their records
// stay in L2) instead of every bra's whole ket range.  Measured on (H2O)_64
// 6-31G: faster for the FP64 lists, slightly slower for the FP32 ones (whose
// chunks are longer and whose Coulomb atomics then meet on the same kets).
bool ket_order_of(const rchem_ctx* ctx, const ClassTune& t) {
  if (ctx->chunk_order_env >= 0) return ctx->chunk_order_env != 0;
  if (t.ket_order >= 0) return t.ket_order != 0;
  return t.fp32 == 0;
}

bool tune_valid(int variant, int blocks, int chunk_tiles) {
  return variant >= 0 && variant < kNumLdsVariants && blocks >= 1 && blocks <= 65536 &&
         chunk_tiles >= 1 && chunk_tiles <= 4096;
}

void set_tune_in(std::vector<ClassTune>* table, const ClassTune& t) {
  for (ClassTune& x : *table)
    if (x.la == t.la && x.lb == t.lb && x.lc == t.lc && x.ld == t.ld && x.fp32 == t.fp32) {
      x = t;
      return;
    }
  table->push_back(t);
}

void set_tune(rchem_ctx* ctx, const ClassTune& t) { set_tune_in(&ctx->tuning, t); }

// Reads a tuning table: one line per (class, precision),
//   la lb lc ld fp32 variant blocks chunk_tiles [ket_order]
// with '#' starting a comment; lines that do not parse or are out of range are
// skipped.  Returns the number of entries taken.
int load_tuning_into(std::vector<ClassTune>* table, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return 0;
  int taken = 0;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    if (char* h = std::strchr(line, '#')) *h = '\0';
    ClassTune t;
    const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %d", &t.la, &t.lb, &t.lc, &t.ld,
                                &t.fp32, &t.variant, &t.blocks, &t.chunk_tiles, &t.ket_order);
    if (got != 8 && got != 9) continue;
    if (got == 9 && t.ket_order != 0 && t.ket_order != 1) continue;
    if (t.la < 0 || t.la > kMaxL || t.lb < 0 || t.lb > kMaxL || t.lc < 0 || t.lc > kMaxL ||
        t.ld < 0 || t.ld > kMaxL || (t.fp32 != 0 && t.fp32 != 1) ||
        !tune_valid(t.variant, t.blocks, t.chunk_tiles))
      continue;
    set_tune_in(table, t);
    ++taken;
  }
  std::fclose(f);
  return taken;
}

int load_tuning(rchem_ctx* ctx, const std::string& path) {
  return load_tuning_into(&ctx->tuning, path);
}

// Where one of the library's own tuning tables lies: data/<file> beside the
// shared library (the environment variable env overrides; empty: no table).
std::string tuning_path(const char* env, const char* file) {
  if (const char* e = std::getenv(env)) return e;
  Dl_info info{};
  if (dladdr(reinterpret_cast<void*>(&load_tuning_into), &info) == 0 || !info.dli_fname) return "";
  std::string so = info.dli_fname;
  const size_t slash = so.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
  return dir + "/data/" + file;
}

int set_device(rchem_ctx* ctx) {
  HIPC(hipSetDevice(ctx->device));
  return RCHEM_OK;
}

// Finishes a basis: function offsets, radial norms folded into coefficients.
int finish_basis(rchem_ctx* ctx) {
  // a rejected basis leaves the context without one (never half of the old one)
  ctx->have_basis = false;
  int f = 0;
  ctx->ccoef.resize(ctx->coefs.size());
  for (auto& s : ctx->shells) {
    for (int x = 0; x < 3; ++x)
      if (!std::isfinite(s.xyz[x])) return ctx->fail(RCHEM_ERR_INVALID, "non-finite centre");
    if (s.l < 0 || s.l > kMaxL)
      return ctx->fail(RCHEM_ERR_UNSUPPORTED,
                       "angular momentum " + std::to_string(s.l) + " not supported (max 2)");
    if (s.nprim <= 0) return ctx->fail(RCHEM_ERR_INVALID, "shell with no primitives");
    if (s.nprim > kMaxPrimPerShell)
      return ctx->fail(RCHEM_ERR_UNSUPPORTED, "more than 8 primitives in a shell");
    s.f0 = f;
    f += ncart(s.l);
    for (int p = 0; p < s.nprim; ++p) {
      const double a = ctx->exps[s.p0 + p];
      if (!(a > 0.0) || !std::isfinite(a))
        return ctx->fail(RCHEM_ERR_INVALID, "non-positive or non-finite exponent");
      if (!std::isfinite(ctx->coefs[s.p0 + p]))
        return ctx->fail(RCHEM_ERR_INVALID, "non-finite contraction coefficient");
      // exponent-dependent part of PGTO::normalization (basis.rs:140-149)
      const int L = s.l;
      const double nrad = std::sqrt(std::pow(2.0, 2.0 * L + 1.5) * std::pow(a, L + 1.5) /
                                    std::pow(M_PI, 1.5));
      ctx->ccoef[s.p0 + p] = ctx->coefs[s.p0 + p] * nrad;
    }
  }
  ctx->nbf = f;
  ctx->have_basis = true;
  ctx->have_schwarz = false;
  ctx->pairs_tau = -1.0;
  ctx->tiles_tau = -1.0;
  ctx->plan_tau = -1.0;
  ctx->all_pairs.clear();
  return RCHEM_OK;
}

// Primitive pairs of shell pair (a, b) whose |K_ab| reaches eps (all of them
// when eps <= 0), in (i of a, j of b) order.
void make_prims(const rchem_ctx* ctx, int a, int b, double eps, std::vector<PrimRec>* out,
                std::vector<double>* alpha = nullptr) {
  const Shell& sa = ctx->shells[a];
  const Shell& sb = ctx->shells[b];
  double AB2 = 0.0;
  for (int x = 0; x < 3; ++x) AB2 += (sa.xyz[x] - sb.xyz[x]) * (sa.xyz[x] - sb.xyz[x]);
  out->clear();
  if (alpha) alpha->clear();
  for (int ip = 0; ip < sa.nprim; ++ip)
    for (int jp = 0; jp < sb.nprim; ++jp) {
      const double al = ctx->exps[sa.p0 + ip], be = ctx->exps[sb.p0 + jp];
      const double z = al + be;
      PrimRec pr;
      pr.zeta = z;
      // P - A = (alpha A + beta B) / zeta - A = -(beta / zeta) (A - B)
      pr.pax = -be / z * (sa.xyz[0] - sb.xyz[0]);
      pr.pay = -be / z * (sa.xyz[1] - sb.xyz[1]);
      pr.paz = -be / z * (sa.xyz[2] - sb.xyz[2]);
      pr.kab = ctx->ccoef[sa.p0 + ip] * ctx->ccoef[sb.p0 + jp] * std::sqrt(2.0) *
               std::pow(M_PI, 1.25) / z * std::exp(-al * be / z * AB2);
      pr.oo2z = 0.5 / z;
      if (eps > 0.0 && std::fabs(pr.kab) < eps) continue;
      out->push_back(pr);
      if (alpha) alpha->push_back(al);
    }
}

// Lays out the selected pairs -- sorted by (class, primitive-pair count,
// Schwarz bound descending) -- with their primitive pairs transposed per
// bucket, uploads both, and records the bounds in layout order in ctx->pairQ.
// Primitive pairs with |K_ab| < prim_eps are dropped (prim_eps <= 0 keeps all);
// a pair left with none is dropped.
int layout_pairs(rchem_ctx* ctx, const std::vector<PairHost>& sel, double prim_eps) {
  std::vector<PairHost> v;
  v.reserve(sel.size());
  std::vector<PrimRec> tmp;
  for (const PairHost& p : sel) {
    make_prims(ctx, p.a, p.b, prim_eps, &tmp);
    if (tmp.empty()) continue;
    PairHost q = p;
    q.np = static_cast<int>(tmp.size());
    v.push_back(q);
  }
  const bool multi = !ctx->shell_mol.empty();
  auto mol_of = [&](const PairHost& p) { return multi ? ctx->shell_mol[p.a] : 0; };
  std::stable_sort(v.begin(), v.end(), [&](const PairHost& x, const PairHost& y) {
    if (x.cls != y.cls) return x.cls < y.cls;
    if (x.np != y.np) return x.np > y.np;
    if (mol_of(x) != mol_of(y)) return mol_of(x) < mol_of(y);
    return x.Q > y.Q;
  });
  ctx->pair_mol.resize(v.size());
  for (size_t k = 0; k < v.size(); ++k) ctx->pair_mol[k] = mol_of(v[k]);
  ctx->pairs.assign(v.size(), PairRec{});
  ctx->pairQ.resize(v.size());
  ctx->pair_shells.resize(v.size());
  std::vector<PrimRec> prims;
  std::vector<double> alphas, tmp_alpha;
  for (int c = 0; c < kNumPairClasses; ++c) ctx->buckets[c].clear();
  size_t i = 0;
  for (int c = 0; c <= kNumPairClasses; ++c) {
    while (i < v.size() && v[i].cls < c) ++i;
    ctx->class_begin[c] = static_cast<int>(i);
  }
  size_t pos = 0;
  while (pos < v.size()) {
    size_t end = pos;
    while (end < v.size() && v[end].cls == v[pos].cls && v[end].np == v[pos].np) ++end;
    const int n = static_cast<int>(end - pos), np = v[pos].np;
    ctx->buckets[v[pos].cls].push_back(Bucket{static_cast<int>(pos), static_cast<int>(end), np});
    const size_t base = prims.size();
    prims.resize(base + static_cast<size_t>(n) * np);
    alphas.resize(base + static_cast<size_t>(n) * np);
    for (int q = 0; q < n; ++q) {
      const PairHost& ph = v[pos + q];
      const Shell& sa = ctx->shells[ph.a];
      const Shell& sb = ctx->shells[ph.b];
      PairRec& r = ctx->pairs[pos + q];
      for (int x = 0; x < 3; ++x) {
        r.A[x] = sa.xyz[x];
        r.AB[x] = sa.xyz[x] - sb.xyz[x];
      }
      r.fa = sa.f0;
      r.fb = sb.f0;
      r.np = np;
      r.p0 = static_cast<int>(base) + q;
      r.ps = n;
      r.flags = (ph.a != ph.b) ? 1 : 0;
      ctx->pairQ[pos + q] = ph.Q;
      ctx->pair_shells[pos + q] = std::make_pair(ph.a, ph.b);
      make_prims(ctx, ph.a, ph.b, prim_eps, &tmp, &tmp_alpha);
      for (int k = 0; k < np; ++k) {
        prims[base + static_cast<size_t>(k) * n + q] = tmp[k];
        alphas[base + static_cast<size_t>(k) * n + q] = tmp_alpha[k];
      }
    }
    pos = end;
  }
  ctx->nprims = prims.size();
  // per bucket, each molecule's pairs [begin, end) (contiguous: sorted by molecule)
  ctx->bucket_segs.assign(kNumPairClasses, {});
  if (multi) {
    const int nm = ctx->nmol;
    for (int c = 0; c < kNumPairClasses; ++c) {
      auto& sg = ctx->bucket_segs[c];
      sg.assign(ctx->buckets[c].size() * static_cast<size_t>(nm), int2{0, 0});
      for (size_t b = 0; b < ctx->buckets[c].size(); ++b) {
        const Bucket& bk = ctx->buckets[c][b];
        for (int m = 0; m < nm; ++m) sg[b * nm + m] = int2{bk.begin, bk.begin};
        int k = bk.begin;
        while (k < bk.end) {
          const int m = ctx->pair_mol[k];
          int e = k;
          while (e < bk.end && ctx->pair_mol[e] == m) ++e;
          sg[b * nm + m] = int2{k, e};
          k = e;
        }
      }
    }
    HIPC(ctx->pair_mol_d.ensure(sizeof(int) * std::max<size_t>(ctx->pair_mol.size(), 1)));
    if (!ctx->pair_mol.empty())
      HIPC(hipMemcpyAsync(ctx->pair_mol_d.p, ctx->pair_mol.data(),
                          sizeof(int) * ctx->pair_mol.size(), hipMemcpyHostToDevice, ctx->stream));
    // each molecule's functions [f_lo, f_hi): the window its K rows need
    std::vector<int> f_lo(nm, ctx->nbf), f_hi(nm, 0);
    for (size_t sidx = 0; sidx < ctx->shells.size(); ++sidx) {
      const int m = ctx->shell_mol[sidx], f = ctx->shells[sidx].f0;
      f_lo[m] = std::min(f_lo[m], f);
      f_hi[m] = std::max(f_hi[m], f + ncart(ctx->shells[sidx].l));
    }
    ctx->mol_kw = 1;
    for (int m = 0; m < nm; ++m) ctx->mol_kw = std::max(ctx->mol_kw, f_hi[m] - f_lo[m]);
    std::vector<int> kc0(ctx->pair_mol.size());
    for (size_t k = 0; k < kc0.size(); ++k) {
      // the window must fit inside [0, nbf): shift it left for the last molecules
      kc0[k] = std::min(f_lo[ctx->pair_mol[k]], ctx->nbf - ctx->mol_kw);
    }
    HIPC(ctx->pair_kc0_d.ensure(sizeof(int) * std::max<size_t>(kc0.size(), 1)));
    if (!kc0.empty())
      HIPC(hipMemcpyAsync(ctx->pair_kc0_d.p, kc0.data(), sizeof(int) * kc0.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));  // kc0 is a local
  }
  HIPC(ctx->alpha_d.ensure(sizeof(double) * std::max<size_t>(alphas.size(), 1)));
  if (!alphas.empty())
    HIPC(hipMemcpyAsync(ctx->alpha_d.p, alphas.data(), sizeof(double) * alphas.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPC(ctx->pairs_d.ensure(sizeof(PairRec) * std::max<size_t>(ctx->pairs.size(), 1)));
  HIPC(ctx->prims_d.ensure(sizeof(PrimRec) * std::max<size_t>(prims.size(), 1)));
  if (!ctx->pairs.empty())
    HIPC(hipMemcpyAsync(ctx->pairs_d.p, ctx->pairs.data(), sizeof(PairRec) * ctx->pairs.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  if (!prims.empty())
    HIPC(hipMemcpyAsync(ctx->prims_d.p, prims.data(), sizeof(PrimRec) * prims.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  std::vector<PrimRecF> primsf(prims.size());
  for (size_t k = 0; k < prims.size(); ++k) {
    const PrimRec& r = prims[k];
    primsf[k] = PrimRecF{static_cast<float>(r.zeta), static_cast<float>(r.pax),
                         static_cast<float>(r.pay), static_cast<float>(r.paz),
                         static_cast<float>(r.kab), static_cast<float>(r.oo2z)};
  }
  HIPC(ctx->prims_f_d.ensure(sizeof(PrimRecF) * std::max<size_t>(primsf.size(), 1)));
  if (!primsf.empty())
    HIPC(hipMemcpyAsync(ctx->prims_f_d.p, primsf.data(), sizeof(PrimRecF) * primsf.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));  // primsf is a local
  return RCHEM_OK;
}

// The kernel that evaluates one class in one mode: the class's straight-line
// per-lane kernel when the library carries it, else the wave-cooperative
// kernel with the class's operation tables (always FP64; it also serves the
// FP32 tile list of a mixed build).
struct ClassKernel {
  EriLaunchFn fn = nullptr;
  EriStreamFn sfn = nullptr;  // streaming kernel (uses gp's HRR and output tables)
  const GenericProgram* gp = nullptr;
  int mode = 0, bm = 0;
  size_t lds_limit = 65536;  // the device's LDS per workgroup (wave-cooperative kernel)
  int shape = 0;             // wave-cooperative kernel's launch shape (0: chosen from the LDS)
  bool valid() const { return fn || gp; }
  hipError_t operator()(const EriArgs& a, int blocks, hipStream_t s) const {
    if (fn) return fn(a, blocks, s);
    if (sfn) return sfn(a, *gp, blocks, s, lds_limit);
    return eri_generic_launch(a, *gp, mode, bm, blocks, s, lds_limit, shape);
  }
};

int upload_programs(rchem_ctx* ctx) {
  if (!ctx->gprog.empty()) return RCHEM_OK;
  auto al = [](size_t b) { return (b + 15) & ~size_t(15); };
  auto room = [](int n, size_t w) { return w * static_cast<size_t>(std::max(n, 1)); };
  size_t total = 0;
  for (int i = 0; i < kNumHostPrograms; ++i) {
    const HostProgram& h = kHostPrograms[i];
    total += al(room(h.nvops, 16)) + al(room(h.nvgroups, 16)) + al(room(h.naccs, 8)) +
             al(room(h.nhops, 8)) + al(room(h.nhgroups, 16)) + al(room(h.nout, 16)) +
             al(room(h.nout, 8));
  }
  std::vector<unsigned char> host(total, 0);
  HIPC(ctx->gprog_d.ensure(total));
  std::vector<GenericProgram> progs(kNumHostPrograms);
  size_t off = 0;
  auto put = [&](const void* src, int n, size_t w) {
    std::memcpy(host.data() + off, src, w * static_cast<size_t>(n));
    const void* dev = static_cast<const unsigned char*>(ctx->gprog_d.p) + off;
    off += al(room(n, w));
    return dev;
  };
  for (int i = 0; i < kNumHostPrograms; ++i) {
    const HostProgram& h = kHostPrograms[i];
    GenericProgram& g = progs[i];
    g.la = h.la; g.lb = h.lb; g.lc = h.lc; g.ld = h.ld;
    g.L = h.L; g.nacc = h.nacc; g.nv = h.nv; g.nh = h.nh; g.nout = h.nout;
    g.nvops = h.nvops; g.nvgroups = h.nvgroups; g.nhops = h.nhops; g.nhgroups = h.nhgroups;
    g.naccs = h.naccs;
    g.P = h.P;
    g.centre = h.grad >= 2 ? h.grad - 2 : -1;
    g.vops = static_cast<const uint4*>(put(h.vops, h.nvops, 16));
    g.vgroups = static_cast<const uint4*>(put(h.vgroups, h.nvgroups, 16));
    g.accs = static_cast<const uint2*>(put(h.accs, h.naccs, 8));
    g.hops = static_cast<const uint2*>(put(h.hops, h.nhops, 8));
    g.hgroups = static_cast<const uint4*>(put(h.hgroups, h.nhgroups, 16));
    g.outs = static_cast<const int4*>(put(h.outs, h.nout, 16));
    g.out_norm = static_cast<const double*>(put(h.out_norm, h.nout, 8));
  }
  HIPC(hipMemcpy(ctx->gprog_d.p, host.data(), total, hipMemcpyHostToDevice));
  ctx->gprog = std::move(progs);
  return RCHEM_OK;
}

// The wave-cooperative program of class (la lb|lc ld): integrals (grad = 0),
// derivative integrals (grad = 1) or those of one centre (grad = 2 + centre);
// nullptr when the library has none.
int find_program(rchem_ctx* ctx, int la, int lb, int lc, int ld, int grad,
                 const GenericProgram** out) {
  *out = nullptr;
  TRY(upload_programs(ctx));
  for (int i = 0; i < kNumHostPrograms; ++i) {
    const HostProgram& h = kHostPrograms[i];
    if (h.la == la && h.lb == lb && h.lc == lc && h.ld == ld && h.grad == grad) {
      if (eri_generic_lds_bytes(ctx->gprog[i]) <= ctx->lds_block) *out = &ctx->gprog[i];
      return RCHEM_OK;
    }
  }
  return RCHEM_OK;
}

// Kernel for class (la lb|lc ld) in `mode`; `out` stays invalid when neither
// engine serves the class (angular momentum beyond the Boys grid's orders).
// Whether the class runs on the streaming kernel (no per-lane accumulators;
// it takes the HRR and output tables from the class's wave-cooperative
// program) where it is the faster engine, measured on the 300-shell s/p/d
// microbenchmark (ms per FP64 build, streaming vs per-lane or
// wave-cooperative): (dp|dp) 138 vs 246, (dd|pp) 78 vs 150, (dd|dp) 84 vs 470,
// (dd|dd) 37 vs 113; the per-lane kernels keep (dp|pp) 125 vs 174, (dp|ds) 52
// vs 142, (dd|ps) 17 vs 72, (dd|ds) 53 vs 62.  RCHEM_ERI_STREAM=0: never, 2:
// every class it serves (also in place of LDS digestion).
bool stream_preferred(const rchem_ctx* ctx, int la, int lb, int lc, int ld) {
  const int key = ((la * 10 + lb) * 10 + lc) * 10 + ld;
  return ctx->eri_stream_mode == 2 ||
         (ctx->eri_stream_mode == 1 && (key == 2121 || key == 2211 || key == 2221 || key == 2222));
}

int class_kernel(rchem_ctx* ctx, int la, int lb, int lc, int ld, int mode, int bm, int fp32,
                 ClassKernel* out) {
  *out = ClassKernel{};
  out->mode = mode;
  out->bm = bm;
  out->lds_limit = ctx->lds_block;
  if (la + lb + lc + ld > kBoysMaxM) return RCHEM_OK;
  if (!ctx->force_generic) {
    const bool prefer = stream_preferred(ctx, la, lb, lc, ld);
    const EriStreamFn sfn =
        prefer && bm == kBoysExact ? eri_stream_launcher(la, lb, lc, ld, mode) : nullptr;
    if (sfn) {
      TRY(find_program(ctx, la, lb, lc, ld, 0, &out->gp));
      if (out->gp) {
        out->sfn = sfn;
        return RCHEM_OK;
      }
    }
    out->fn = eri_launcher(la, lb, lc, ld, mode, bm, fp32);
    if (out->fn) return RCHEM_OK;
  }
  return find_program(ctx, la, lb, lc, ld, 0, &out->gp);
}

std::string class_name(int la, int lb, int lc, int ld) {
  return "(" + std::to_string(la) + std::to_string(lb) + "|" + std::to_string(lc) +
         std::to_string(ld) + ")";
}

int upload_tiles(rchem_ctx* ctx) {
  HIPC(ctx->tiles_d.ensure(sizeof(Tile) * std::max<size_t>(ctx->tiles.size(), 1)));
  if (!ctx->tiles.empty())
    HIPC(hipMemcpyAsync(ctx->tiles_d.p, ctx->tiles.data(), sizeof(Tile) * ctx->tiles.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return RCHEM_OK;
}

void class_ls(int X, int* l1, int* l2) {
  for (int a = 0; a <= kMaxL; ++a)
    for (int b = 0; b <= a; ++b)
      if (pair_class(a, b) == X) {
        *l1 = a;
        *l2 = b;
      }
}

// Schwarz bounds for every shell pair (exact Boys), once per basis.
int compute_schwarz(rchem_ctx* ctx) {
  if (ctx->have_schwarz) return RCHEM_OK;
  const int ns = static_cast<int>(ctx->shells.size());
  std::vector<PairHost> all;
  all.reserve(static_cast<size_t>(ns) * (ns + 1) / 2);
  const bool multi = !ctx->shell_mol.empty();
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) {
      if (multi && ctx->shell_mol[a] != ctx->shell_mol[b]) continue;
      int x = a, y = b;
      if (ctx->shells[y].l > ctx->shells[x].l) std::swap(x, y);
      all.push_back(PairHost{x, y, pair_class(ctx->shells[x].l, ctx->shells[y].l),
                             ctx->shells[x].nprim * ctx->shells[y].nprim, 0.0});
    }
  TRY(layout_pairs(ctx, all, 0.0));
  // diagonal tiles: 64 consecutive pairs of one bucket, bra = ket per lane
  ctx->tiles.clear();
  std::vector<std::pair<int, int>> ranges;  // per class: tile range
  for (int X = 0; X < kNumPairClasses; ++X) {
    const int t0 = static_cast<int>(ctx->tiles.size());
    for (const Bucket& bk : ctx->buckets[X])
      for (int j = bk.begin; j < bk.end; j += 64)
        ctx->tiles.push_back(Tile{-1, j, std::min(64, bk.end - j), 0});
    ranges.emplace_back(t0, static_cast<int>(ctx->tiles.size()) - t0);
  }
  TRY(upload_tiles(ctx));
  DevBuf q;
  HIPC(q.ensure(sizeof(double) * std::max<size_t>(ctx->pairs.size(), 1)));
  for (int X = 0; X < kNumPairClasses; ++X) {
    if (ranges[X].second == 0) continue;
    int l1, l2;
    class_ls(X, &l1, &l2);
    ClassKernel fn;
    TRY(class_kernel(ctx, l1, l2, l1, l2, kModeDiag, kBoysExact, 0, &fn));
    if (!fn.valid())
      return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " + class_name(l1, l2, l1, l2) +
                                                  " not served by this library");
    EriArgs a{};
    a.pairs = ctx->pairs_d.as<PairRec>();
    a.prims = ctx->prims_d.as<PrimRec>();
  a.prims_f = ctx->prims_f_d.as<PrimRecF>();
    a.tiles = ctx->tiles_d.as<Tile>() + ranges[X].first;
    a.ntiles = ranges[X].second;
    a.nbf = ctx->nbf;
    a.boys_table = ctx->boys_d.as<double>();
    a.boys_table_f = ctx->boys_f_d.as<float>();
    a.out = q.as<double>();
    HIPC(fn(a, 0, ctx->stream));
  }
  std::vector<double> Q(ctx->pairs.size());
  if (!Q.empty())
    HIPC(hipMemcpyAsync(Q.data(), q.p, sizeof(double) * Q.size(), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  // the records are in layout order; ctx->pair_shells names each one's shells
  std::vector<PairHost> sorted(ctx->pairs.size());
  for (size_t i = 0; i < sorted.size(); ++i) {
    if (!std::isfinite(Q[i]))
      return ctx->fail(RCHEM_ERR_NONFINITE, "non-finite Schwarz bound");
    const int a = ctx->pair_shells[i].first, b = ctx->pair_shells[i].second;
    sorted[i] = PairHost{a, b, pair_class(ctx->shells[a].l, ctx->shells[b].l),
                         ctx->shells[a].nprim * ctx->shells[b].nprim, Q[i]};
  }
  ctx->all_pairs = std::move(sorted);
  ctx->have_schwarz = true;
  ctx->pairs_tau = -1.0;
  ctx->tiles_tau = -1.0;
  ctx->plan_tau = -1.0;
  return RCHEM_OK;
}

int flops_for(int la, int lb, int lc, int ld, int* vrr, int* hrr) {
  for (int r = 0; r < kEriFlopsRows; ++r)
    if (kEriFlops[r][0] == la && kEriFlops[r][1] == lb && kEriFlops[r][2] == lc &&
        kEriFlops[r][3] == ld) {
      *vrr = kEriFlops[r][4];
      *hrr = kEriFlops[r][5];
      return 1;
    }
  return 0;
}

// Pair list (Schwarz- and primitive-screened) for threshold tau.
int prepare_pairs(rchem_ctx* ctx, double tau) {
  TRY(compute_schwarz(ctx));
  if (ctx->pairs_tau == tau) return RCHEM_OK;
  double qmax = 0.0;
  for (const auto& p : ctx->all_pairs) qmax = std::max(qmax, p.Q);
  std::vector<PairHost> sel;
  sel.reserve(ctx->all_pairs.size());
  for (const auto& p : ctx->all_pairs)
    if (tau <= 0.0 || p.Q * qmax >= tau) sel.push_back(p);
  // A primitive pair enters an integral through K_ab K_cd / sqrt(zeta + eta)
  // times factors of order one; dropping |K_ab| < 1e-3 tau / K_max keeps every
  // dropped primitive quartet three orders below the quartet threshold.
  double kmax = 0.0;
  if (tau > 0.0) {
    std::vector<PrimRec> tmp;
    for (const auto& p : sel) {
      make_prims(ctx, p.a, p.b, 0.0, &tmp);
      for (const auto& pr : tmp) kmax = std::max(kmax, std::fabs(pr.kab));
    }
  }
  const double prim_eps = (tau > 0.0 && kmax > 0.0) ? 1e-3 * tau / kmax : 0.0;
  TRY(layout_pairs(ctx, sel, prim_eps));
  ctx->pairs_tau = tau;
  ctx->tiles_tau = -1.0;
  ctx->plan_tau = -1.0;
  return RCHEM_OK;
}

// Walks every (quartet class, bra pair, ket bucket) task of threshold tau and
// calls f(class, bra, bucket, ket_begin, ket_end) for its Schwarz survivors.
template <typename Fn>
void for_each_task(const rchem_ctx* ctx, double tau, Fn f) {
  const std::vector<double>& Q = ctx->pairQ;
  for (int X = 0; X < kNumPairClasses; ++X)
    for (int Y = 0; Y <= X; ++Y)
      for (int i = ctx->class_begin[X]; i < ctx->class_begin[X + 1]; ++i)
        for (size_t b = 0; b < ctx->buckets[Y].size(); ++b) {
          const Bucket& bk = ctx->buckets[Y][b];
          int lo = bk.begin, hi = bk.end;
          if (!ctx->shell_mol.empty()) {  // the bra's molecule only
            const int2 sg = ctx->bucket_segs[Y][b * ctx->nmol + ctx->pair_mol[i]];
            lo = sg.x;
            hi = sg.y;
          }
          if (X == Y) hi = (lo > i) ? lo : std::min(hi, i + 1);
          if (tau > 0.0) {
            if (Q[i] <= 0.0) {
              hi = lo;
            } else {
              const double need = tau / Q[i];
              int l = lo, h = hi;
              while (l < h) {
                const int m = (l + h) / 2;
                if (Q[m] >= need) l = m + 1; else h = m;
              }
              hi = l;
            }
          }
          f(X, Y, i, static_cast<int>(b), lo, hi);
        }
}

QClass make_qclass(int X, int Y) {
  QClass qc{};
  qc.X = X;
  qc.Y = Y;
  class_ls(X, &qc.la, &qc.lb);
  class_ls(Y, &qc.lc, &qc.ld);
  return qc;
}

double per_prim_flops(const QClass& qc, int* hrr) {
  int fv = 0, fh = 0;
  flops_for(qc.la, qc.lb, qc.lc, qc.ld, &fv, &fh);
  *hrr = fh;
  return fv + 68.0 + 4.0 * (qc.la + qc.lb + qc.lc + qc.ld);
}

// Host-built tile list (used for the dense tensor, where tau = 0 and systems
// are small).
int prepare_host_tiles(rchem_ctx* ctx, double tau) {
  TRY(prepare_pairs(ctx, tau));
  if (ctx->tiles_tau == tau) return RCHEM_OK;
  ctx->tiles.clear();
  ctx->hclasses.clear();
  int curX = -1, curY = -1;
  QClass qc{};
  auto flush = [&]() {
    qc.ntiles = static_cast<int>(ctx->tiles.size()) - qc.tile_begin;
    if (curX >= 0 && qc.ntiles > 0) ctx->hclasses.push_back(qc);
  };
  for_each_task(ctx, tau, [&](int X, int Y, int i, int, int lo, int hi) {
    if (X != curX || Y != curY) {
      flush();
      curX = X;
      curY = Y;
      qc = make_qclass(X, Y);
      qc.tile_begin = static_cast<int>(ctx->tiles.size());
    }
    for (int j = lo; j < hi; j += 64) ctx->tiles.push_back(Tile{i, j, std::min(64, hi - j), 0});
  });
  flush();
  for (const QClass& q : ctx->hclasses) {
    ClassKernel fn;
    TRY(class_kernel(ctx, q.la, q.lb, q.lc, q.ld, kModeTensor, kBoysExact, 0, &fn));
    if (!fn.valid())
      return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " + class_name(q.la, q.lb, q.lc, q.ld) +
                                                  " not served by this library");
  }
  TRY(upload_tiles(ctx));
  ctx->tiles_tau = tau;
  return RCHEM_OK;
}

// Device-planner tables for threshold tau: classes, buckets, capacities and
// the FP64 statistics (exact for all-FP64 builds).
constexpr double kPassTiles = 67108864.0;  // 2^26 tiles (1 GiB per tile list) per pass

int prepare_plan(rchem_ctx* ctx, double tau) {
  TRY(prepare_pairs(ctx, tau));
  if (ctx->plan_tau == tau) return RCHEM_OK;
  ctx->pclasses.clear();
  ctx->seg_class.clear();
  ctx->seg_cap.clear();
  ctx->passes.clear();
  ctx->qclasses.clear();
  std::vector<int4> bks;
  std::vector<int2> segs;
  std::vector<int> cls_bk_off;
  for (int X = 0; X < kNumPairClasses; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      const int nbra = ctx->class_begin[X + 1] - ctx->class_begin[X];
      const int nbk = static_cast<int>(ctx->buckets[Y].size());
      if (nbra == 0 || nbk == 0) continue;
      cls_bk_off.push_back(static_cast<int>(bks.size()));
      for (const Bucket& bk : ctx->buckets[Y]) bks.push_back(make_int4(bk.begin, bk.end, bk.np, 0));
      if (!ctx->shell_mol.empty())
        segs.insert(segs.end(), ctx->bucket_segs[Y].begin(), ctx->bucket_segs[Y].end());
      ctx->qclasses.push_back(make_qclass(X, Y));
    }
  // tiles reserved per bra pair (one partial tile more per task covers the
  // FP64/FP32 split) and the FP64 statistics
  const size_t nq = ctx->qclasses.size();
  std::vector<std::vector<double>> bra_cap(nq);
  rchem_jk_stats st{};
  std::vector<int> cls_index(kNumPairClasses * kNumPairClasses, -1);
  for (size_t c = 0; c < nq; ++c) {
    const QClass& qc = ctx->qclasses[c];
    cls_index[qc.X * kNumPairClasses + qc.Y] = static_cast<int>(c);
    bra_cap[c].assign(ctx->class_begin[qc.X + 1] - ctx->class_begin[qc.X], 0.0);
  }
  for (auto& qc : ctx->qclasses) qc.quartets = qc.prim_quartets = qc.flops = 0.0;
  for_each_task(ctx, tau, [&](int X, int Y, int i, int b, int lo, int hi) {
    const int c = cls_index[X * kNumPairClasses + Y];
    if (c < 0) return;
    const int cnt = hi - lo;
    bra_cap[c][i - ctx->class_begin[X]] += (cnt + 63) / 64 + 1;
    if (cnt <= 0) return;
    QClass& qc = ctx->qclasses[c];
    const double pq = static_cast<double>(cnt) * ctx->pairs[i].np * ctx->buckets[Y][b].np;
    qc.quartets += cnt;
    qc.prim_quartets += pq;
  });
  for (size_t c = 0; c < nq; ++c) {
    QClass& qc = ctx->qclasses[c];
    int fh = 0;
    const double fp = per_prim_flops(qc, &fh);
    qc.flops = qc.prim_quartets * fp + qc.quartets * fh;
    st.quartets += qc.quartets;
    st.primitive_quartets += qc.prim_quartets;
    st.integrals += qc.quartets * ncart(qc.la) * ncart(qc.lb) * ncart(qc.lc) * ncart(qc.ld);
    st.flops += qc.flops;
    if (qc.quartets > 0) {
      ClassKernel fn;
      TRY(class_kernel(ctx, qc.la, qc.lb, qc.lc, qc.ld, kModeJK, kBoysExact, 0, &fn));
      if (!fn.valid())
        return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " +
                                                    class_name(qc.la, qc.lb, qc.lc, qc.ld) +
                                                    " not served by this library");
    }
  }
  // segments and passes
  rchem_ctx::Pass cur{0, 0, 0, 0, 0, 0, 0};
  double pass_cap = 0.0, pass_ccap = 0.0;
  auto close_pass = [&]() {
    if (cur.nseg == 0) return;
    cur.cap = static_cast<int>(pass_cap) + 1;
    cur.ccap = static_cast<int>(std::min(pass_ccap, pass_cap)) + 1;
    ctx->passes.push_back(cur);
    cur = rchem_ctx::Pass{static_cast<int>(ctx->pclasses.size()), 0, 0, 0, 0, 0, 0};
    pass_cap = 0.0;
    pass_ccap = 0.0;
  };
  auto add_segment = [&](int c, int lo, int hi, double cap) {
    const QClass& qc = ctx->qclasses[c];
    PlanClass pc{};
    pc.bra_begin = lo;
    pc.bra_end = hi;
    pc.bk_off = cls_bk_off[c];
    pc.nbk = static_cast<int>(ctx->buckets[qc.Y].size());
    pc.task_begin = cur.tasks;
    pc.same = qc.X == qc.Y;
    pc.unit_begin = cur.units;
    const ClassTune t64 = tune_for(ctx, qc.la, qc.lb, qc.lc, qc.ld, 0);
    const ClassTune t32 = tune_for(ctx, qc.la, qc.lb, qc.lc, qc.ld, 1);
    pc.chunk64 = t64.chunk_tiles;
    pc.chunk32 = t32.chunk_tiles;
    pc.order = (ket_order_of(ctx, t64) ? 1 : 0) | (ket_order_of(ctx, t32) ? 2 : 0);
    cur.order |= pc.order;
    ctx->pclasses.push_back(pc);
    ctx->seg_class.push_back(c);
    ctx->seg_cap.push_back(static_cast<int>(cap));
    cur.tasks += (hi - lo) * pc.nbk;
    cur.units += hi - lo;
    cur.nseg += 1;
    pass_cap += cap;
    // a unit's chunks: ceil(tiles / chunk size) <= tiles / chunk size + 1
    pass_ccap += cap / std::max(1, std::min(pc.chunk64, pc.chunk32)) + (hi - lo);
  };
  for (size_t c = 0; c < nq; ++c) {
    const int b0 = ctx->class_begin[ctx->qclasses[c].X];
    const int nb = static_cast<int>(bra_cap[c].size());
    int lo = 0;
    double acc = 0.0;
    for (int i = 0; i < nb; ++i) {
      const double bc = bra_cap[c][i];
      if (bc > 2.0e9) return ctx->fail(RCHEM_ERR_ALLOC, "quartet tile list of one bra pair too large");
      if (pass_cap + acc + bc > kPassTiles && (acc > 0.0 || pass_cap > 0.0)) {
        if (acc > 0.0) add_segment(static_cast<int>(c), b0 + lo, b0 + i, acc);
        close_pass();
        lo = i;
        acc = 0.0;
      }
      acc += bc;
    }
    if (nb > lo) add_segment(static_cast<int>(c), b0 + lo, b0 + nb, acc);
  }
  close_pass();
  int tasks = 0, units = 0;
  ctx->plan_cap = 1;
  for (const auto& ps : ctx->passes) {
    tasks = std::max(tasks, ps.tasks);
    units = std::max(units, ps.units);
    ctx->plan_cap = std::max(ctx->plan_cap, ps.cap);
  }
  ctx->plan_tasks = tasks;
  ctx->plan_units = units;
  const size_t nt = static_cast<size_t>(tasks) + 1;
  const size_t nu = static_cast<size_t>(units) + 1;
  for (DevBuf* b : {&ctx->plan_chc64, &ctx->plan_chc32, &ctx->plan_choff64, &ctx->plan_choff32})
    HIPC(b->ensure(sizeof(int) * nu));
  HIPC(ctx->plan_chunks64.ensure(sizeof(Chunk) * static_cast<size_t>(ctx->plan_cap)));
  HIPC(ctx->plan_chunks32.ensure(sizeof(Chunk) * static_cast<size_t>(ctx->plan_cap)));
  ctx->plan_key_bits = 0;
  ctx->plan_ccap = 0;
  bool any_order = false;
  for (const auto& ps : ctx->passes) any_order = any_order || ps.order != 0;
  if (any_order) {
    int nseg = 1;
    for (const auto& ps : ctx->passes) {
      nseg = std::max(nseg, ps.nseg);
      ctx->plan_ccap = std::max(ctx->plan_ccap, ps.ccap);
    }
    int ket_bits = 0;
    while ((static_cast<size_t>(1) << ket_bits) <= (ctx->pairQ.size() >> 6)) ++ket_bits;
    const int bits = plan_class_bits(nseg) + ket_bits;
    if (bits <= 32 && ctx->plan_ccap > 0) {
      ctx->plan_key_bits = std::max(bits, 1);
      const size_t cc = static_cast<size_t>(ctx->plan_ccap);
      HIPC(ctx->plan_chunks_raw.ensure(sizeof(Chunk) * 2 * cc));
      HIPC(ctx->plan_keys_raw.ensure(sizeof(unsigned) * 2 * cc));
      HIPC(ctx->plan_keys_sorted.ensure(sizeof(unsigned) * cc));
      ctx->plan_sort_bytes = plan_sort_bytes(ctx->plan_ccap);
      HIPC(ctx->plan_sort_tmp.ensure(std::max<size_t>(ctx->plan_sort_bytes, 8)));
    }
  }
  // pair-indexed Coulomb accumulators, one block per pair class
  ctx->jp_cls.assign(kNumPairClasses, JpClass{});
  std::vector<int> ncomp(2 * kNumPairClasses);
  size_t jp = 0;
  for (int X = 0; X < kNumPairClasses; ++X) {
    int l1, l2;
    class_ls(X, &l1, &l2);
    ncomp[2 * X] = ncart(l1);
    ncomp[2 * X + 1] = ncart(l2);
    const int nX = ctx->class_begin[X + 1] - ctx->class_begin[X];
    ctx->jp_cls[X] = JpClass{static_cast<int>(jp), nX, ctx->class_begin[X], 0};
    jp += static_cast<size_t>(ncart(l1) * ncart(l2)) * nX;
  }
  ctx->jp_size = jp;
  HIPC(ctx->jp_d.ensure(sizeof(double) * std::max<size_t>(jp, 1)));
  HIPC(ctx->jp_cls_d.ensure(sizeof(JpClass) * kNumPairClasses));
  HIPC(ctx->jp_ncomp_d.ensure(sizeof(int) * ncomp.size()));
  HIPC(hipMemcpyAsync(ctx->jp_cls_d.p, ctx->jp_cls.data(), sizeof(JpClass) * kNumPairClasses,
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->jp_ncomp_d.p, ncomp.data(), sizeof(int) * ncomp.size(),
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(ctx->plan_classes_d.ensure(sizeof(PlanClass) * std::max<size_t>(ctx->pclasses.size(), 1)));
  HIPC(ctx->plan_buckets_d.ensure(sizeof(int4) * std::max<size_t>(bks.size(), 1)));
  HIPC(ctx->plan_q_d.ensure(sizeof(double) * std::max<size_t>(ctx->pairQ.size(), 1)));
  for (DevBuf* b : {&ctx->plan_cnt64, &ctx->plan_cnt32, &ctx->plan_off64, &ctx->plan_off32,
                    &ctx->plan_lo, &ctx->plan_mid, &ctx->plan_hi})
    HIPC(b->ensure(sizeof(int) * nt));
  HIPC(ctx->plan_tiles64.ensure(sizeof(Tile) * static_cast<size_t>(ctx->plan_cap)));
  HIPC(ctx->plan_tiles32.ensure(sizeof(Tile) * static_cast<size_t>(ctx->plan_cap)));
  HIPC(ctx->plan_ranges.ensure(sizeof(int) * 8 * std::max<size_t>(ctx->pclasses.size(), 1)));
  HIPC(ctx->plan_stats.ensure(sizeof(double) * 4 * kPlanStatsReplicas *
                              std::max<size_t>(ctx->pclasses.size(), 1)));
  HIPC(ctx->plan_dmax.ensure(sizeof(double)));
  ctx->plan_scan_bytes = std::max(plan_scan_bytes(tasks), plan_scan_bytes(units));
  HIPC(ctx->plan_scan_tmp.ensure(std::max<size_t>(ctx->plan_scan_bytes, 8)));
  if (!ctx->pclasses.empty())
    HIPC(hipMemcpyAsync(ctx->plan_classes_d.p, ctx->pclasses.data(),
                        sizeof(PlanClass) * ctx->pclasses.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  if (!bks.empty())
    HIPC(hipMemcpyAsync(ctx->plan_buckets_d.p, bks.data(), sizeof(int4) * bks.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  if (!segs.empty()) {
    HIPC(ctx->plan_segs_d.ensure(sizeof(int2) * segs.size()));
    HIPC(hipMemcpyAsync(ctx->plan_segs_d.p, segs.data(), sizeof(int2) * segs.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  }
  if (!ctx->pairQ.empty())
    HIPC(hipMemcpyAsync(ctx->plan_q_d.p, ctx->pairQ.data(), sizeof(double) * ctx->pairQ.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->planned = st;
  ctx->plan_tau = tau;
  return RCHEM_OK;
}

int check_opts(rchem_ctx* ctx, const rchem_opts* o) {
  if (o->boys_mode != RCHEM_BOYS_EXACT && o->boys_mode != RCHEM_BOYS_COMPAT)
    return ctx->fail(RCHEM_ERR_INVALID, "unknown boys_mode");
  if (o->precision != RCHEM_PREC_FP64 && o->precision != RCHEM_PREC_MIXED)
    return ctx->fail(RCHEM_ERR_INVALID, "unknown precision");
  if (o->precision == RCHEM_PREC_MIXED && o->boys_mode != RCHEM_BOYS_EXACT)
    return ctx->fail(RCHEM_ERR_UNSUPPORTED, "mixed precision needs the exact Boys function");
  if (o->precision == RCHEM_PREC_MIXED && !(o->mixed_thresh >= 0.0))
    return ctx->fail(RCHEM_ERR_INVALID, "mixed_thresh must be >= 0");
  if (!(o->schwarz_thresh >= 0.0) || !std::isfinite(o->schwarz_thresh))
    return ctx->fail(RCHEM_ERR_INVALID, "schwarz_thresh must be finite and >= 0");
  return RCHEM_OK;
}

int check_scf_opts(rchem_ctx* ctx, const rchem_scf_opts* o) {
  if (!(o->damping >= 0.0 && o->damping < 1.0))
    return ctx->fail(RCHEM_ERR_INVALID, "damping must lie in [0, 1)");
  if (!(o->damping_until >= 0.0) || !std::isfinite(o->damping_until))
    return ctx->fail(RCHEM_ERR_INVALID, "damping_until must be finite and >= 0");
  return RCHEM_OK;
}

ClassTimer* timer_for(rchem_ctx* ctx, const QClass& qc, int fp32) {
  for (auto& t : ctx->timers)
    if (t.la == qc.la && t.lb == qc.lb && t.lc == qc.lc && t.ld == qc.ld && t.fp32 == fp32)
      return &t;
  ClassTimer t;
  t.la = qc.la; t.lb = qc.lb; t.lc = qc.lc; t.ld = qc.ld;
  t.fp32 = fp32;
  t.cls = static_cast<int>(&qc - ctx->qclasses.data());
  ctx->timers.push_back(t);
  return &ctx->timers.back();
}

int take_events(rchem_ctx* ctx, std::pair<hipEvent_t, hipEvent_t>* ev) {
  if (!ctx->event_pool.empty()) {
    *ev = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return RCHEM_OK;
  }
  HIPC(hipEventCreate(&ev->first));
  HIPC(hipEventCreate(&ev->second));
  return RCHEM_OK;
}

// Folds completed event pairs into the per-class sums (synchronises on them).
int collect_timers(rchem_ctx* ctx) {
  for (auto& t : ctx->timers) {
    for (auto& ev : t.pending) {
      HIPC(hipEventSynchronize(ev.second));
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, ev.first, ev.second));
      t.ms += ms;
      t.launches += 1;
      ctx->event_pool.push_back(ev);
    }
    t.pending.clear();
  }
  return RCHEM_OK;
}

EriArgs base_args(const rchem_ctx* ctx) {
  EriArgs a{};
  a.pairs = ctx->pairs_d.as<PairRec>();
  a.prims = ctx->prims_d.as<PrimRec>();
  a.prims_f = ctx->prims_f_d.as<PrimRecF>();
  a.nbf = ctx->nbf;
  a.boys_table = ctx->boys_d.as<double>();
  a.boys_table_f = ctx->boys_f_d.as<float>();
  return a;
}

// Host-tiled launch of every class in `mode` (diagonal / tensor modes).
int run_eri_host(rchem_ctx* ctx, int mode, int bm, double* out) {
  for (const QClass& qc : ctx->hclasses) {
    ClassKernel fn;
    TRY(class_kernel(ctx, qc.la, qc.lb, qc.lc, qc.ld, mode, bm, 0, &fn));
    if (!fn.valid())
      return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " + class_name(qc.la, qc.lb, qc.lc, qc.ld) +
                                                  " not served by this library");
    EriArgs a = base_args(ctx);
    a.tiles = ctx->tiles_d.as<Tile>() + qc.tile_begin;
    a.ntiles = qc.ntiles;
    a.out = out;
    HIPC(fn(a, 0, ctx->stream));
  }
  return RCHEM_OK;
}

// LDS an LDS-digestion workgroup needs beside the K rows (its Boys block and
// exp row: at most 23.1 KB in FP64)
constexpr size_t kLdsStaticReserve = 24 * 1024;

int launch_timed(rchem_ctx* ctx, const QClass& qc, int fp32, const ClassKernel& fn,
                 const EriArgs& a, int blocks, bool may_time = true) {
  const bool timed = ctx->profiling && may_time;
  std::pair<hipEvent_t, hipEvent_t> ev{};
  if (timed) {
    TRY(take_events(ctx, &ev));
    HIPC(hipEventRecord(ev.first, ctx->stream));
  }
  HIPC(fn(a, blocks, ctx->stream));
  if (timed) {
    HIPC(hipEventRecord(ev.second, ctx->stream));
    ClassTimer* t = timer_for(ctx, qc, fp32);
    t->quartets = qc.quartets;
    t->prim_quartets = qc.prim_quartets;
    t->flops = qc.flops;
    t->pending.push_back(ev);
  }
  return RCHEM_OK;
}

// Per-build accumulators of a Fock build: the lower-half J and K, the
// pair-indexed Coulomb accumulators and density blocks, the planner's
// per-class statistics and max |D|.  Every one is cleared or fully rewritten
// at the start of a build.
struct JkBuffers {
  double* Jh;
  double* Kh;
  double* jp;
  double* dp;
  double* stats;
  double* dmax;
};

// Doubles a build's JkBuffers take for the current plan, each part rounded up
// to 256 bytes.
size_t jk_buffer_doubles(const rchem_ctx* ctx) {
  auto up = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
  const size_t jp = std::max<size_t>(ctx->jp_size, 1);
  const size_t nc = std::max<size_t>(ctx->pclasses.size(), 1);
  return 2 * up(n2) + 2 * up(jp) + up(4 * nc * kPlanStatsReplicas) + up(1);
}

// The build's accumulators carved out of one device range of
// jk_buffer_doubles(ctx) doubles.
JkBuffers jk_buffers_in(const rchem_ctx* ctx, double* base) {
  auto up = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
  const size_t jp = std::max<size_t>(ctx->jp_size, 1);
  const size_t nc = std::max<size_t>(ctx->pclasses.size(), 1);
  JkBuffers b{};
  double* p = base;
  b.Jh = p;
  p += up(n2);
  b.Kh = p;
  p += up(n2);
  b.jp = p;
  p += up(jp);
  b.dp = p;
  p += up(jp);
  b.stats = p;
  p += up(4 * nc * kPlanStatsReplicas);
  b.dmax = p;
  return b;
}

// One Fock build.  With ext == nullptr the accumulators are the context's own
// (allocated on first use) and per-class timing follows ctx->profiling; with
// ext the build allocates nothing and records no events (the plan for the
// threshold must exist already).
int jk_device(rchem_ctx* ctx, const double* dD, double* dJ, double* dK, const rchem_opts* o,
              const JkBuffers* ext = nullptr) {
  TRY(prepare_plan(ctx, o->schwarz_thresh));
  const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
  const bool mixed = o->precision == RCHEM_PREC_MIXED;
  ctx->last_mixed_pending = mixed;
  JkBuffers bf{};
  if (ext) {
    bf = *ext;
  } else {
    HIPC(ctx->work_Jh.ensure(n2 * sizeof(double)));
    HIPC(ctx->work_Kh.ensure(n2 * sizeof(double)));
    HIPC(ctx->dp_d.ensure(sizeof(double) * std::max<size_t>(ctx->jp_size, 1)));
    bf = JkBuffers{ctx->work_Jh.as<double>(), ctx->work_Kh.as<double>(), ctx->jp_d.as<double>(),
                   ctx->dp_d.as<double>(), ctx->plan_stats.as<double>(),
                   ctx->plan_dmax.as<double>()};
  }
  ctx->last_stats_dev = bf.stats;
  const bool timed_ok = ext == nullptr;
  const bool blocks = ctx->nmol > 1 && ctx->mol_nmax > 0;  // block-diagonal J and K
  if (blocks) {
    HIPC(fill_blocks_launch(bf.Jh, bf.Kh, ctx->nbf, ctx->mol_f0_d.as<int>(), ctx->mol_nb_d.as<int>(),
                            ctx->nmol, ctx->mol_nmax, ctx->stream));
  } else {
    HIPC(fill_launch(bf.Jh, n2, 0.0, ctx->stream));
    HIPC(fill_launch(bf.Kh, n2, 0.0, ctx->stream));
  }
  const int nc = static_cast<int>(ctx->pclasses.size());
  HIPC(fill_launch(bf.stats, 4 * kPlanStatsReplicas * static_cast<size_t>(std::max(nc, 1)), 0.0,
                   ctx->stream));
  if (mixed) {
    HIPC(fill_launch(bf.dmax, 1, 0.0, ctx->stream));
    if (ctx->nmol > 1 && ctx->mol_nmax > 0)  // a multi-molecule basis: its diagonal blocks
      HIPC(dmax_blocks_launch(dD, static_cast<size_t>(ctx->nbf), ctx->mol_f0_d.as<int>(),
                              ctx->mol_nb_d.as<int>(), ctx->nmol, ctx->mol_nmax,
                              reinterpret_cast<unsigned long long*>(bf.dmax), ctx->stream));
    else
      HIPC(dmax_launch(dD, n2, reinterpret_cast<unsigned long long*>(bf.dmax), ctx->stream));
  }
  HIPC(fill_launch(bf.jp, std::max<size_t>(ctx->jp_size, 1), 0.0, ctx->stream));
  // the pairs' density blocks in the Jp layout, for the LDS kernels' Coulomb sums
  HIPC(dpair_gather_launch(ctx->pairs_d.as<PairRec>(), dD, ctx->jp_cls_d.as<JpClass>(),
                           ctx->jp_ncomp_d.as<int>(), kNumPairClasses,
                           static_cast<int>(ctx->pairs.size()), bf.dp, ctx->nbf,
                           ctx->stream));
  for (const auto& ps : ctx->passes) {
    PlanArgs pa{};
    pa.classes = ctx->plan_classes_d.as<PlanClass>() + ps.seg_begin;
    pa.nclasses = ps.nseg;
    pa.ntasks = ps.tasks;
    pa.nunits = ps.units;
    pa.buckets = ctx->plan_buckets_d.as<int4>();
    pa.Q = ctx->plan_q_d.as<double>();
    pa.pairs = ctx->pairs_d.as<PairRec>();
    pa.tau = o->schwarz_thresh;
    pa.mixed_thresh = mixed ? o->mixed_thresh : 0.0;
    pa.dmax = bf.dmax;
    pa.cnt64 = ctx->plan_cnt64.as<int>();
    pa.cnt32 = ctx->plan_cnt32.as<int>();
    pa.lo = ctx->plan_lo.as<int>();
    pa.mid = ctx->plan_mid.as<int>();
    pa.hi = ctx->plan_hi.as<int>();
    pa.stats = bf.stats + 4 * static_cast<size_t>(ps.seg_begin);
    pa.stats_stride = 4 * std::max(nc, 1);
    if (!ctx->shell_mol.empty()) {
      pa.segs = ctx->plan_segs_d.as<int2>();
      pa.nmol = ctx->nmol;
      pa.pair_mol = ctx->pair_mol_d.as<int>();
      if (ctx->mol_active_on) pa.mol_active = ctx->mol_active_d.as<int>();
    }
    PlanBuffers pb{};
    pb.off64 = ctx->plan_off64.as<int>();
    pb.off32 = ctx->plan_off32.as<int>();
    pb.tiles64 = ctx->plan_tiles64.as<Tile>();
    pb.tiles32 = ctx->plan_tiles32.as<Tile>();
    pb.chc64 = ctx->plan_chc64.as<int>();
    pb.chc32 = ctx->plan_chc32.as<int>();
    pb.choff64 = ctx->plan_choff64.as<int>();
    pb.choff32 = ctx->plan_choff32.as<int>();
    pb.chunks64 = ctx->plan_chunks64.as<Chunk>();
    pb.chunks32 = ctx->plan_chunks32.as<Chunk>();
    pb.cap = ps.cap;
    pb.ranges = ctx->plan_ranges.as<int>() + 8 * static_cast<size_t>(ps.seg_begin);
    pb.scan_tmp = ctx->plan_scan_tmp.p;
    pb.scan_bytes = ctx->plan_scan_bytes;
    if (ctx->plan_key_bits > 0 && ps.order != 0) {
      pb.key_bits = ctx->plan_key_bits;
      pb.sort_mask = ps.order;
      pb.ccap = std::min(ps.ccap, ctx->plan_ccap);
      pb.chunks_raw = ctx->plan_chunks_raw.as<Chunk>();
      pb.keys_raw = ctx->plan_keys_raw.as<unsigned>();
      pb.keys_sorted = ctx->plan_keys_sorted.as<unsigned>();
      pb.sort_tmp = ctx->plan_sort_tmp.p;
      pb.sort_bytes = ctx->plan_sort_bytes;
    }
    HIPC(plan_launch(pa, pb, ctx->stream));
    for (int sg = ps.seg_begin; sg < ps.seg_begin + ps.nseg; ++sg) {
      const int c = ctx->seg_class[sg];
      const QClass& qc = ctx->qclasses[c];
      if (qc.quartets <= 0) continue;
      const int cap = ctx->seg_cap[sg];
      EriArgs a = base_args(ctx);
      a.D = dD;
      a.J = bf.Jh;
      a.K = bf.Kh;
      a.Jp = bf.jp;
      a.jp_bra = ctx->jp_cls[qc.X];
      a.jp_ket = ctx->jp_cls[qc.Y];
      a.Dp = bf.dp;
      // K rows of the bra pair in LDS when they fit in 64 KB: every column, or
      // the bra's molecule's window in a multi-molecule basis
      const bool multi = !ctx->shell_mol.empty();
      a.kw = multi ? ctx->mol_kw : ctx->nbf;
      a.pair_kc0 = multi ? ctx->pair_kc0_d.as<int>() : nullptr;
      // K rows in LDS when they fit the device's LDS per workgroup beside the
      // kernel's own arrays (Boys block, bra primitive pairs: < 12 KB)
      const size_t lds = sizeof(double) * (ncart(qc.la) + ncart(qc.lb)) * static_cast<size_t>(a.kw);
      const bool lds_ok = lds + kLdsStaticReserve <= ctx->lds_block && ctx->lds_digestion &&
                          !ctx->force_generic &&
                          !(qc.la + qc.lb + qc.lc + qc.ld >= 5 &&
                            stream_preferred(ctx, qc.la, qc.lb, qc.lc, qc.ld));
      int* ranges = pb.ranges + 8 * (sg - ps.seg_begin);
      for (int prec = 0; prec < (mixed ? 2 : 1); ++prec) {
        const int bm = prec ? kBoysExact : o->boys_mode;
        a.tiles = prec ? ctx->plan_tiles32.as<Tile>() : ctx->plan_tiles64.as<Tile>();
        const ClassTune tn = tune_for(ctx, qc.la, qc.lb, qc.lc, qc.ld, prec);
        ClassKernel fn;
        if (lds_ok) {
          fn.fn = eri_jk_lds_launcher(qc.la, qc.lb, qc.lc, qc.ld, bm, prec, tn.variant);
          if (!fn.fn && tn.variant != 0)  // shape not compiled for this class
            fn.fn = eri_jk_lds_launcher(qc.la, qc.lb, qc.lc, qc.ld, bm, prec, 0);
        }
        if (fn.valid()) {
          a.chunks = prec ? ctx->plan_chunks32.as<Chunk>() : ctx->plan_chunks64.as<Chunk>();
          a.chunk_range = ranges + (prec ? 6 : 4);
          a.tile_range = nullptr;
          TRY(launch_timed(ctx, qc, prec, fn, a, std::min(tn.blocks, cap), timed_ok));
        } else {
          TRY(class_kernel(ctx, qc.la, qc.lb, qc.lc, qc.ld, kModeJK, bm, prec, &fn));
          if (!fn.valid())
            return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " +
                                                        class_name(qc.la, qc.lb, qc.lc, qc.ld) +
                                                        " not served by this library");
          fn.shape = tn.variant;  // the class's launch shape on the wave-cooperative kernel
          a.tile_range = ranges + (prec ? 2 : 0);
          TRY(launch_timed(ctx, qc, prec, fn, a, std::min(tn.blocks, (cap + 3) / 4), timed_ok));
        }
      }
    }
    if (ctx->progress && !ctx->capturing) {
      // a long build (many passes) reports each finished pass
      HIPC(hipStreamSynchronize(ctx->stream));
      std::fprintf(stderr, "[rchem] Fock build: pass %d of %zu done\n",
                   static_cast<int>(&ps - ctx->passes.data()) + 1, ctx->passes.size());
      std::fflush(stderr);
    }
  }
  HIPC(jp_scatter_launch(ctx->pairs_d.as<PairRec>(), bf.jp, ctx->jp_cls_d.as<JpClass>(),
                         ctx->jp_ncomp_d.as<int>(), kNumPairClasses,
                         static_cast<int>(ctx->pairs.size()), bf.Jh, ctx->nbf, ctx->stream));
  if (blocks)
    HIPC(symmetrize_blocks_launch(bf.Jh, bf.Kh, dJ, dK, ctx->nbf, ctx->mol_f0_d.as<int>(),
                                  ctx->mol_nb_d.as<int>(), ctx->nmol, ctx->mol_nmax,
                                  ctx->stream));
  else
    HIPC(symmetrize_launch(bf.Jh, bf.Kh, dJ, dK, ctx->nbf, ctx->stream));
  ctx->last_stats = ctx->planned;
  ctx->last_mixed = mixed;
  return RCHEM_OK;
}

// Device memory of rchem_jk_device_noalloc: the accumulators of one Fock
// build, in a file-scope array every GPU holds from module load to process end
// (nothing allocated per call).  Host code never passes its name to a kernel:
// in host code it stands for a host-side shadow, so the device address of each
// GPU's copy is looked up once per device (hipGetSymbolAddress, with that
// device current) when the first context on the device is created, and only
// that address is used afterwards.  Calls on one GPU share the array.
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

struct NoAllocScratch {
  std::mutex mu;
  double* ptr[kNoAllocMaxDevices] = {};
  hipError_t status[kNoAllocMaxDevices] = {};
  bool tried[kNoAllocMaxDevices] = {};
};
NoAllocScratch& noalloc_scratch() {
  static NoAllocScratch s;
  return s;
}

// Resolves the current device's copy of `scratch` (once per device; the
// caller has made `device` current).
void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return;
  NoAllocScratch& s = noalloc_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  if (s.tried[device]) return;
  s.tried[device] = true;
  void* p = nullptr;
  s.status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (s.status[device] == hipSuccess) s.ptr[device] = static_cast<double*>(p);
}

double* noalloc_scratch_ptr(int device, hipError_t* status) {
  *status = hipErrorInvalidDevice;
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  NoAllocScratch& s = noalloc_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  *status = s.tried[device] ? s.status[device] : hipErrorNotInitialized;
  return s.ptr[device];
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int rchem_opts_default(rchem_opts* o) {
  if (!o) return RCHEM_ERR_INVALID;
  o->boys_mode = RCHEM_BOYS_EXACT;
  o->precision = RCHEM_PREC_FP64;
  o->schwarz_thresh = 1e-12;
  o->mixed_thresh = 1e-7;
  return RCHEM_OK;
}

int rchem_ctx_create(int device, rchem_ctx** out) {
  if (!out) return RCHEM_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_create_error = "no HIP device visible";
    return RCHEM_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "device index out of range";
    return RCHEM_ERR_INVALID;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    g_create_error = "hipGetDeviceProperties failed";
    return RCHEM_ERR_HIP;
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_create_error = std::string("device is ") + prop.gcnArchName + ", library built for gfx950";
    return RCHEM_ERR_NO_DEVICE;
  }
  auto* ctx = new rchem_ctx();
  ctx->device = device;
  ctx->lds_block = prop.sharedMemPerBlock;
  {
    // the Fock build's table, and the batched SCF's (multi-molecule bases of
    // small molecules: a bra pair has few kets, so other shapes win)
    const std::string tp = tuning_path("RCHEM_TUNING_FILE", "tuning_gfx950.txt");
    if (!tp.empty()) load_tuning(ctx, tp);
    const std::string bp = tuning_path("RCHEM_TUNING_BATCH_FILE", "tuning_gfx950_batch.txt");
    if (!bp.empty()) load_tuning_into(&ctx->tuning_batch, bp);
  }
  if (const char* e = std::getenv("RCHEM_LDS_DIGESTION")) ctx->lds_digestion = std::atoi(e) != 0;
  if (const char* e = std::getenv("RCHEM_ERI_GENERIC")) ctx->force_generic = std::atoi(e) != 0;
  if (const char* e = std::getenv("RCHEM_ERI_STREAM")) ctx->eri_stream_mode = std::atoi(e);
  if (const char* e = std::getenv("RCHEM_PROGRESS")) ctx->progress = std::atoi(e) != 0;
  if (const char* e = std::getenv("RCHEM_CHUNK_ORDER")) ctx->chunk_order_env = std::atoi(e) != 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    g_create_error = "stream creation failed";
    delete ctx;
    return RCHEM_ERR_HIP;
  }
  if (rocblas_create_handle(&ctx->blas) != rocblas_status_success ||
      rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success) {
    g_create_error = "rocblas handle creation failed";
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RCHEM_ERR_HIP;
  }
  noalloc_scratch_init(device);  // device is current (hipSetDevice above)
  std::vector<double> tab = make_boys_table();
  std::vector<float> tabf(tab.begin(), tab.end());
  if (ctx->boys_d.ensure(tab.size() * sizeof(double)) != hipSuccess ||
      hipMemcpy(ctx->boys_d.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess ||
      ctx->boys_f_d.ensure(tabf.size() * sizeof(float)) != hipSuccess ||
      hipMemcpy(ctx->boys_f_d.p, tabf.data(), tabf.size() * sizeof(float),
                hipMemcpyHostToDevice) != hipSuccess) {
    g_create_error = "Boys table upload failed";
    rocblas_destroy_handle(ctx->blas);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RCHEM_ERR_HIP;
  }
  *out = ctx;
  return RCHEM_OK;
}

int rchem_ctx_destroy(rchem_ctx* ctx) {
  if (!ctx) return RCHEM_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->timers)
    for (auto& ev : t.pending) ctx->event_pool.push_back(ev);
  for (auto& ev : ctx->event_pool) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  if (ctx->blas) rocblas_destroy_handle(ctx->blas);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->scf_status_h) (void)hipHostFree(ctx->scf_status_h);
  delete ctx;
  return RCHEM_OK;
}

const char* rchem_last_error(const rchem_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int rchem_basis_upload(rchem_ctx* ctx, int nshell, const double* xyz, const int* l,
                       const int* nprim, const double* exps, const double* coefs) {
  if (ctx) {
    ctx->shell_mol.clear();
    ctx->nmol = 1;
    ctx->mol_nmax = 0;
  }
  if (!ctx) return RCHEM_ERR_INVALID;
  if (nshell <= 0 || !xyz || !l || !nprim || !exps || !coefs)
    return ctx->fail(RCHEM_ERR_INVALID, "rchem_basis_upload: bad arguments");
  ctx->have_basis = false;
  ctx->shells.assign(nshell, Shell{});
  ctx->exps.clear();
  ctx->coefs.clear();
  for (int s = 0; s < nshell; ++s) {
    Shell& sh = ctx->shells[s];
    sh.l = l[s];
    sh.nprim = nprim[s];
    if (sh.nprim <= 0) return ctx->fail(RCHEM_ERR_INVALID, "shell with no primitives");
    sh.p0 = static_cast<int>(ctx->exps.size());
    for (int x = 0; x < 3; ++x) {
      sh.xyz[x] = xyz[3 * s + x];
      if (!std::isfinite(sh.xyz[x])) return ctx->fail(RCHEM_ERR_INVALID, "non-finite centre");
    }
    for (int p = 0; p < sh.nprim; ++p) {
      ctx->exps.push_back(exps[sh.p0 + p]);
      ctx->coefs.push_back(coefs[sh.p0 + p]);
    }
  }
  TRY(set_device(ctx));
  return finish_basis(ctx);
}

int rchem_basis_from_atoms(rchem_ctx* ctx, int natom, const int* Z, const double* xyz,
                           const char* basis_name) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (natom <= 0 || !Z || !xyz || !basis_name)
    return ctx->fail(RCHEM_ERR_INVALID, "rchem_basis_from_atoms: bad arguments");
  auto lower = [](std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
  };
  const std::string want = lower(basis_name);
  bool known = false;
  for (int r = 0; r < kBasisNumRows; ++r)
    if (lower(kBasisRows[r].set) == want) known = true;
  if (!known) return ctx->fail(RCHEM_ERR_INVALID, std::string("unknown basis set ") + basis_name);
  ctx->have_basis = false;
  ctx->shells.clear();
  ctx->exps.clear();
  ctx->coefs.clear();
  ctx->shell_mol.clear();
  ctx->nmol = 1;
  ctx->mol_nmax = 0;
  // per atom, per basis-set shell, per l in its angular_momentum list (basis.rs:186-205)
  for (int a = 0; a < natom; ++a) {
    bool found = false;
    for (int r = 0; r < kBasisNumRows; ++r) {
      const BasisShellRow& row = kBasisRows[r];
      if (lower(row.set) != want || row.Z != Z[a]) continue;
      found = true;
      for (int k = 0; k < 2; ++k) {
        const int l = row.am[k];
        if (l < 0) continue;
        Shell sh;
        sh.l = l;
        sh.nprim = row.nprim;
        sh.p0 = static_cast<int>(ctx->exps.size());
        for (int x = 0; x < 3; ++x) sh.xyz[x] = xyz[3 * a + x];
        for (int p = 0; p < row.nprim; ++p) {
          ctx->exps.push_back(row.exps[p]);
          ctx->coefs.push_back(row.coefs[k][p]);  // coefficients[l] (basis.rs:201)
        }
        ctx->shells.push_back(sh);
      }
    }
    if (!found)
      return ctx->fail(RCHEM_ERR_INVALID, "element Z=" + std::to_string(Z[a]) + " not in " +
                                              basis_name);
  }
  TRY(set_device(ctx));
  return finish_basis(ctx);
}

int rchem_basis_nbf(const rchem_ctx* ctx, int* nbf) {
  if (!ctx || !nbf) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return RCHEM_ERR_NO_BASIS;
  *nbf = ctx->nbf;
  return RCHEM_OK;
}

int rchem_basis_dims(const rchem_ctx* ctx, int* nshell, int* nprim_total) {
  if (!ctx || !nshell || !nprim_total) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return RCHEM_ERR_NO_BASIS;
  *nshell = static_cast<int>(ctx->shells.size());
  *nprim_total = static_cast<int>(ctx->exps.size());
  return RCHEM_OK;
}

int rchem_basis_shells(const rchem_ctx* ctx, double* xyz, int* l, int* nprim, double* exps,
                       double* coefs) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return RCHEM_ERR_NO_BASIS;
  for (size_t s = 0; s < ctx->shells.size(); ++s) {
    const Shell& sh = ctx->shells[s];
    if (xyz)
      for (int x = 0; x < 3; ++x) xyz[3 * s + x] = sh.xyz[x];
    if (l) l[s] = sh.l;
    if (nprim) nprim[s] = sh.nprim;
  }
  if (exps) std::copy(ctx->exps.begin(), ctx->exps.end(), exps);
  if (coefs) std::copy(ctx->coefs.begin(), ctx->coefs.end(), coefs);
  return RCHEM_OK;
}

int rchem_jk_device(rchem_ctx* ctx, const double* dD, double* dJ, double* dK,
                    const rchem_opts* opts) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!dD || !dJ || !dK) return ctx->fail(RCHEM_ERR_INVALID, "null device pointer");
  rchem_opts o;
  rchem_opts_default(&o);
  if (opts) o = *opts;
  TRY(check_opts(ctx, &o));
  TRY(set_device(ctx));
  return jk_device(ctx, dD, dJ, dK, &o);
}

int rchem_jk_device_noalloc(rchem_ctx* ctx, const double* dD, double* dJ, double* dK,
                            const rchem_opts* opts) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!dD || !dJ || !dK) return ctx->fail(RCHEM_ERR_INVALID, "null device pointer");
  rchem_opts o;
  rchem_opts_default(&o);
  if (opts) o = *opts;
  TRY(check_opts(ctx, &o));
  // the plan (pair lists, tile lists, planner tables) belongs to the context
  // and is built by the first rchem_jk / rchem_jk_device call at a threshold;
  // building it allocates, so this entry point only uses an existing one
  if (ctx->plan_tau != o.schwarz_thresh || ctx->pairs_tau != o.schwarz_thresh)
    return ctx->fail(RCHEM_ERR_INVALID,
                     "rchem_jk_device_noalloc: no plan for this Schwarz threshold on the "
                     "context (call rchem_jk_device once with the same options first)");
  const size_t need = jk_buffer_doubles(ctx) * sizeof(double);
  if (need > kNoAllocScratchBytes)
    return ctx->fail(RCHEM_ERR_INVALID,
                     "rchem_jk_device_noalloc: this basis needs " + std::to_string(need >> 20) +
                         " MiB of build accumulators, beyond the fixed " +
                         std::to_string(kNoAllocScratchBytes >> 20) + " MiB");
  hipError_t st = hipSuccess;
  double* base = noalloc_scratch_ptr(ctx->device, &st);
  if (!base)
    return ctx->fail(RCHEM_ERR_HIP, std::string("rchem_jk_device_noalloc: device scratch "
                                                "address unavailable: ") +
                                        hipGetErrorString(st));
  TRY(set_device(ctx));
  const JkBuffers bf = jk_buffers_in(ctx, base);
  return jk_device(ctx, dD, dJ, dK, &o, &bf);
}

int rchem_lds_variant_info(int variant, int* waves, int* min_waves_per_simd, int* prefetch,
                           int* unroll, int* branch_free) {
  if (variant < 0 || variant >= kNumLdsVariants) return RCHEM_ERR_INVALID;
  const LdsVariant& v = kLdsVariants[variant];
  if (waves) *waves = v.waves;
  if (min_waves_per_simd) *min_waves_per_simd = v.min_waves_per_simd;
  if (prefetch) *prefetch = v.prefetch;
  if (unroll) *unroll = v.unroll;
  if (branch_free) *branch_free = v.branch_free;
  return RCHEM_OK;
}

int rchem_set_class_tuning(rchem_ctx* ctx, int la, int lb, int lc, int ld, int fp32, int variant,
                           int blocks, int chunk_tiles) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (la < 0 || la > kMaxL || lb < 0 || lb > kMaxL || lc < 0 || lc > kMaxL || ld < 0 ||
      ld > kMaxL || (fp32 != 0 && fp32 != 1) || !tune_valid(variant, blocks, chunk_tiles))
    return ctx->fail(RCHEM_ERR_INVALID, "rchem_set_class_tuning: argument out of range");
  ClassTune t;
  t.la = la; t.lb = lb; t.lc = lc; t.ld = ld; t.fp32 = fp32;
  t.variant = variant;
  t.blocks = blocks;
  t.chunk_tiles = chunk_tiles;
  set_tune(ctx, t);
  ctx->plan_tau = -1.0;  // chunk sizes are part of the plan
  return RCHEM_OK;
}

int rchem_set_class_ket_order(rchem_ctx* ctx, int la, int lb, int lc, int ld, int fp32,
                              int ket_order) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (la < 0 || la > kMaxL || lb < 0 || lb > kMaxL || lc < 0 || lc > kMaxL || ld < 0 ||
      ld > kMaxL || (fp32 != 0 && fp32 != 1) || ket_order < -1 || ket_order > 1)
    return ctx->fail(RCHEM_ERR_INVALID, "rchem_set_class_ket_order: argument out of range");
  ClassTune t = tune_for(ctx, la, lb, lc, ld, fp32);
  t.ket_order = ket_order;
  set_tune(ctx, t);
  ctx->plan_tau = -1.0;  // the chunk order is part of the plan
  return RCHEM_OK;
}

int rchem_get_class_ket_order(const rchem_ctx* ctx, int la, int lb, int lc, int ld, int fp32,
                              int* ket_order) {
  if (!ctx || !ket_order) return RCHEM_ERR_INVALID;
  *ket_order = ket_order_of(ctx, tune_for(ctx, la, lb, lc, ld, fp32 ? 1 : 0)) ? 1 : 0;
  return RCHEM_OK;
}

int rchem_get_class_tuning(const rchem_ctx* ctx, int la, int lb, int lc, int ld, int fp32,
                           int* variant, int* blocks, int* chunk_tiles) {
  if (!ctx || !variant || !blocks || !chunk_tiles) return RCHEM_ERR_INVALID;
  const ClassTune t = tune_for(ctx, la, lb, lc, ld, fp32 ? 1 : 0);
  *variant = t.variant;
  *blocks = t.blocks;
  *chunk_tiles = t.chunk_tiles;
  return RCHEM_OK;
}

int rchem_load_tuning(rchem_ctx* ctx, const char* path, int* entries) {
  if (!ctx) return RCHEM_ERR_INVALID;
  ctx->tuning.clear();
  ctx->plan_tau = -1.0;
  const int n = (path && *path) ? load_tuning(ctx, path) : 0;
  if (entries) *entries = n;
  return RCHEM_OK;
}

int rchem_load_batch_tuning(rchem_ctx* ctx, const char* path, int* entries) {
  if (!ctx) return RCHEM_ERR_INVALID;
  ctx->tuning_batch.clear();
  const int n = (path && *path) ? load_tuning_into(&ctx->tuning_batch, path) : 0;
  if (entries) *entries = n;
  return RCHEM_OK;
}

int rchem_jk(rchem_ctx* ctx, const double* D, double* J, double* K, const rchem_opts* opts) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!D || !J || !K) return ctx->fail(RCHEM_ERR_INVALID, "null pointer");
  rchem_opts o;
  rchem_opts_default(&o);
  if (opts) o = *opts;
  TRY(check_opts(ctx, &o));
  TRY(set_device(ctx));
  const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
  for (size_t i = 0; i < n2; ++i)
    if (!std::isfinite(D[i])) return ctx->fail(RCHEM_ERR_NONFINITE, "non-finite density");
  HIPC(ctx->work_D.ensure(n2 * sizeof(double)));
  HIPC(ctx->work_J.ensure(n2 * sizeof(double)));
  HIPC(ctx->work_K.ensure(n2 * sizeof(double)));
  HIPC(hipMemcpyAsync(ctx->work_D.p, D, n2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  TRY(jk_device(ctx, ctx->work_D.as<double>(), ctx->work_J.as<double>(), ctx->work_K.as<double>(),
                &o));
  HIPC(hipMemcpyAsync(J, ctx->work_J.p, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipMemcpyAsync(K, ctx->work_K.p, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return RCHEM_OK;
}

int rchem_set_profiling(rchem_ctx* ctx, int enable) {
  if (!ctx) return RCHEM_ERR_INVALID;
  ctx->profiling = enable != 0;
  return RCHEM_OK;
}

// where the last build's per-class planner statistics went
static const double* last_stats_src(const rchem_ctx* ctx) {
  return ctx->last_stats_dev ? ctx->last_stats_dev : ctx->plan_stats.as<double>();
}

// The last build's per-class statistics {quartets64, quartets32, prim64,
// prim32} (4 per segment): the sum of the planner's replicas.
static int read_last_stats(rchem_ctx* ctx, std::vector<double>* out) {
  const size_t per = 4 * ctx->pclasses.size();
  std::vector<double> all(per * kPlanStatsReplicas);
  HIPC(hipMemcpyAsync(all.data(), last_stats_src(ctx), sizeof(double) * all.size(),
                      hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  out->assign(per, 0.0);
  for (size_t r = 0; r < static_cast<size_t>(kPlanStatsReplicas); ++r)
    for (size_t i = 0; i < per; ++i) (*out)[i] += all[r * per + i];
  return RCHEM_OK;
}

int rchem_reset_class_timings(rchem_ctx* ctx) {
  if (!ctx) return RCHEM_ERR_INVALID;
  TRY(set_device(ctx));
  TRY(collect_timers(ctx));
  ctx->timers.clear();
  return RCHEM_OK;
}

int rchem_jk_class_timings(rchem_ctx* ctx, rchem_class_timing* out, int max, int* count) {
  if (!ctx || !count || (max > 0 && !out)) return RCHEM_ERR_INVALID;
  TRY(set_device(ctx));
  TRY(collect_timers(ctx));
  if (ctx->last_mixed && !ctx->pclasses.empty()) {
    // work per precision of the last build, counted by the planner on the device
    std::vector<double> dv;
    TRY(read_last_stats(ctx, &dv));
    for (auto& t : ctx->timers) {
      if (t.cls < 0 || t.cls >= static_cast<int>(ctx->qclasses.size())) continue;
      const QClass& qc = ctx->qclasses[t.cls];
      int fh = 0;
      const double fp = per_prim_flops(qc, &fh);
      t.quartets = t.prim_quartets = 0.0;
      for (size_t sg = 0; sg < ctx->seg_class.size(); ++sg)
        if (ctx->seg_class[sg] == t.cls) {
          t.quartets += dv[4 * sg + (t.fp32 ? 1 : 0)];
          t.prim_quartets += dv[4 * sg + (t.fp32 ? 3 : 2)];
        }
      t.flops = t.prim_quartets * fp + t.quartets * fh;
    }
  }
  *count = static_cast<int>(ctx->timers.size());
  for (int i = 0; i < max && i < *count; ++i) {
    const ClassTimer& t = ctx->timers[i];
    out[i] = rchem_class_timing{t.la, t.lb, t.lc, t.ld, t.fp32, t.launches, t.ms,
                                t.quartets, t.prim_quartets, t.flops};
  }
  return RCHEM_OK;
}

int rchem_jk_last_stats(rchem_ctx* ctx, rchem_jk_stats* st) {
  if (!ctx || !st) return RCHEM_ERR_INVALID;
  *st = ctx->last_stats;
  if (!ctx->pclasses.empty() && ctx->plan_tau >= 0.0) {
    // tiles of the last build: the planner's per-class ranges
    TRY(set_device(ctx));
    std::vector<int> rg(8 * ctx->pclasses.size());
    HIPC(hipMemcpyAsync(rg.data(), ctx->plan_ranges.p, sizeof(int) * rg.size(),
                        hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    double tiles = 0.0;
    for (size_t c = 0; c < ctx->pclasses.size(); ++c)
      tiles += (rg[8 * c + 1] - rg[8 * c]) + (rg[8 * c + 3] - rg[8 * c + 2]);
    st->tiles = tiles;
  }
  if (ctx->last_mixed && !ctx->pclasses.empty()) {
    // the planner counted FP64 and FP32 quartets of the last build on the device
    TRY(set_device(ctx));
    std::vector<double> dv;
    TRY(read_last_stats(ctx, &dv));
    double q32 = 0.0;
    for (size_t c = 0; c < ctx->pclasses.size(); ++c) q32 += dv[4 * c + 1];
    st->quartets_fp32 = q32;
  }
  return RCHEM_OK;
}

int rchem_eri_tensor(rchem_ctx* ctx, double* I, const rchem_opts* opts) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!I) return ctx->fail(RCHEM_ERR_INVALID, "null pointer");
  rchem_opts o;
  rchem_opts_default(&o);
  if (opts) o = *opts;
  TRY(check_opts(ctx, &o));
  TRY(set_device(ctx));
  const size_t n = static_cast<size_t>(ctx->nbf);
  const size_t n4 = n * n * n * n;
  if (n > 200) return ctx->fail(RCHEM_ERR_INVALID, "tensor too large (nbf > 200)");
  TRY(prepare_host_tiles(ctx, 0.0));
  DevBuf t;
  HIPC(t.ensure(n4 * sizeof(double)));
  HIPC(fill_launch(t.as<double>(), n4, 0.0, ctx->stream));
  TRY(run_eri_host(ctx, kModeTensor, o.boys_mode, t.as<double>()));
  HIPC(hipMemcpyAsync(I, t.p, n4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return RCHEM_OK;
}

int rchem_eri_shell_quartets(rchem_ctx* ctx, int nq, const int* shells, double* out,
                             const rchem_opts* opts) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (nq < 0 || (nq > 0 && (!shells || !out)))
    return ctx->fail(RCHEM_ERR_INVALID, "bad quartet list");
  rchem_opts o;
  rchem_opts_default(&o);
  if (opts) o = *opts;
  TRY(check_opts(ctx, &o));
  TRY(set_device(ctx));
  const int ns = static_cast<int>(ctx->shells.size());
  for (int k = 0; k < 4 * nq; ++k)
    if (shells[k] < 0 || shells[k] >= ns) return ctx->fail(RCHEM_ERR_INVALID, "shell index out of range");
  if (nq == 0) return RCHEM_OK;
  {
    size_t total = 0;
    for (int q = 0; q < nq; ++q) {
      size_t blk = 1;
      for (int k = 0; k < 4; ++k) blk *= ncart(ctx->shells[shells[4 * q + k]].l);
      total += blk;
    }
    if (total > static_cast<size_t>(std::numeric_limits<int>::max()))
      return ctx->fail(RCHEM_ERR_INVALID, "quartet list too long: output exceeds 2^31 values");
  }
  // Per quartet: two pair records in canonical orientation (higher l first in
  // a pair, higher pair class as bra), their primitive pairs stored
  // contiguously, and a one-lane tile tagged with the quartet index.
  std::vector<PairRec> prs;
  std::vector<PrimRec> pps, tmp;
  struct Job { int perm[4]; int cls; size_t out_off; };
  std::vector<Job> jobs(nq);
  std::vector<std::vector<Tile>> by_class(kNumPairClasses * kNumPairClasses);
  std::vector<std::pair<int, int>> cls_ls(kNumPairClasses * kNumPairClasses);
  size_t off = 0;
  auto add_pair = [&](int a, int b) {
    PairRec r{};
    const Shell& sa = ctx->shells[a];
    const Shell& sb = ctx->shells[b];
    for (int x = 0; x < 3; ++x) {
      r.A[x] = sa.xyz[x];
      r.AB[x] = sa.xyz[x] - sb.xyz[x];
    }
    r.fa = sa.f0;
    r.fb = sb.f0;
    make_prims(ctx, a, b, 0.0, &tmp);
    r.np = static_cast<int>(tmp.size());
    r.p0 = static_cast<int>(pps.size());
    r.ps = 1;
    r.flags = a != b;
    pps.insert(pps.end(), tmp.begin(), tmp.end());
    prs.push_back(r);
    return static_cast<int>(prs.size()) - 1;
  };
  for (int q = 0; q < nq; ++q) {
    int pos[4] = {0, 1, 2, 3};  // canonical slot k holds requested position pos[k]
    const int* s4 = shells + 4 * q;
    if (ctx->shells[s4[1]].l > ctx->shells[s4[0]].l) std::swap(pos[0], pos[1]);
    if (ctx->shells[s4[3]].l > ctx->shells[s4[2]].l) std::swap(pos[2], pos[3]);
    const int X = pair_class(ctx->shells[s4[pos[0]]].l, ctx->shells[s4[pos[1]]].l);
    const int Y = pair_class(ctx->shells[s4[pos[2]]].l, ctx->shells[s4[pos[3]]].l);
    if (X < Y) {
      std::swap(pos[0], pos[2]);
      std::swap(pos[1], pos[3]);
    }
    Job& j = jobs[q];
    for (int k = 0; k < 4; ++k) j.perm[k] = pos[k];
    const int bx = std::max(X, Y), ky = std::min(X, Y);
    j.cls = bx * kNumPairClasses + ky;
    j.out_off = off;
    size_t blk = 1;
    for (int k = 0; k < 4; ++k) blk *= ncart(ctx->shells[s4[k]].l);
    off += blk;
    const int bra = add_pair(s4[pos[0]], s4[pos[1]]);
    const int ket = add_pair(s4[pos[2]], s4[pos[3]]);
    // the tile's last field is the block's offset in the packed output
    by_class[j.cls].push_back(Tile{bra, ket, 1, static_cast<int>(j.out_off)});
    cls_ls[j.cls] = {bx, ky};
  }
  DevBuf dpr, dpp, dt, dout;
  std::vector<Tile> tiles;
  std::vector<std::pair<int, int>> ranges(by_class.size());
  for (size_t c = 0; c < by_class.size(); ++c) {
    ranges[c] = {static_cast<int>(tiles.size()), static_cast<int>(by_class[c].size())};
    tiles.insert(tiles.end(), by_class[c].begin(), by_class[c].end());
  }
  HIPC(dpr.ensure(sizeof(PairRec) * prs.size()));
  HIPC(dpp.ensure(sizeof(PrimRec) * pps.size()));
  HIPC(dt.ensure(sizeof(Tile) * tiles.size()));
  HIPC(dout.ensure(sizeof(double) * off));
  HIPC(hipMemcpyAsync(dpr.p, prs.data(), sizeof(PairRec) * prs.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(dpp.p, pps.data(), sizeof(PrimRec) * pps.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(dt.p, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  for (size_t c = 0; c < by_class.size(); ++c) {
    if (ranges[c].second == 0) continue;
    int la, lb, lc, ld;
    class_ls(cls_ls[c].first, &la, &lb);
    class_ls(cls_ls[c].second, &lc, &ld);
    ClassKernel fn;
    TRY(class_kernel(ctx, la, lb, lc, ld, kModeQuartet, o.boys_mode, 0, &fn));
    if (!fn.valid())
      return ctx->fail(RCHEM_ERR_UNSUPPORTED, "ERI class " + class_name(la, lb, lc, ld) +
                                                  " not served by this library");
    EriArgs a = base_args(ctx);
    a.pairs = dpr.as<PairRec>();
    a.prims = dpp.as<PrimRec>();
    a.tiles = dt.as<Tile>() + ranges[c].first;
    a.ntiles = ranges[c].second;
    a.out = dout.as<double>();
    HIPC(fn(a, 0, ctx->stream));
  }
  std::vector<double> raw(off);
  HIPC(hipMemcpyAsync(raw.data(), dout.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  // canonical block -> requested order
  for (int q = 0; q < nq; ++q) {
    const int* s4 = shells + 4 * q;
    const Job& j = jobs[q];
    int nreq[4], ncan[4];
    for (int k = 0; k < 4; ++k) nreq[k] = ncart(ctx->shells[s4[k]].l);
    for (int k = 0; k < 4; ++k) ncan[k] = nreq[j.perm[k]];
    int c[4];
    for (c[0] = 0; c[0] < nreq[0]; ++c[0])
      for (c[1] = 0; c[1] < nreq[1]; ++c[1])
        for (c[2] = 0; c[2] < nreq[2]; ++c[2])
          for (c[3] = 0; c[3] < nreq[3]; ++c[3]) {
            const size_t src = ((static_cast<size_t>(c[j.perm[0]]) * ncan[1] + c[j.perm[1]]) *
                                    ncan[2] + c[j.perm[2]]) * ncan[3] + c[j.perm[3]];
            const size_t dst = ((static_cast<size_t>(c[0]) * nreq[1] + c[1]) * nreq[2] + c[2]) *
                                   nreq[3] + c[3];
            out[j.out_off + dst] = raw[j.out_off + src];
          }
  }
  return RCHEM_OK;
}

int rchem_schwarz(rchem_ctx* ctx, double* Q, const rchem_opts* opts) {
  (void)opts;
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!Q) return ctx->fail(RCHEM_ERR_INVALID, "null pointer");
  TRY(set_device(ctx));
  TRY(compute_schwarz(ctx));
  const size_t ns = ctx->shells.size();
  std::fill(Q, Q + ns * ns, 0.0);
  for (const auto& p : ctx->all_pairs) {
    const int a = std::max(p.a, p.b), b = std::min(p.a, p.b);
    Q[a * ns + b] = p.Q;
  }
  return RCHEM_OK;
}

int rchem_one_electron(rchem_ctx* ctx, int natom, const double* Z, const double* xyz, double* S,
                       double* T, double* V) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (natom < 0 || (natom > 0 && (!Z || !xyz)))
    return ctx->fail(RCHEM_ERR_INVALID, "bad nuclei");
  for (int a = 0; a < natom; ++a)
    if (!std::isfinite(Z[a]) || !std::isfinite(xyz[3 * a]) || !std::isfinite(xyz[3 * a + 1]) ||
        !std::isfinite(xyz[3 * a + 2]))
      return ctx->fail(RCHEM_ERR_INVALID, "non-finite nuclear charge or position");
  TRY(set_device(ctx));
  const int ns = static_cast<int>(ctx->shells.size());
  std::vector<ShellDev> sh(ns);
  for (int s = 0; s < ns; ++s) {
    for (int x = 0; x < 3; ++x) sh[s].xyz[x] = ctx->shells[s].xyz[x];
    sh[s].l = ctx->shells[s].l;
    sh[s].nprim = ctx->shells[s].nprim;
    sh[s].p0 = ctx->shells[s].p0;
    sh[s].f0 = ctx->shells[s].f0;
  }
  std::vector<int2> pr;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) pr.push_back(make_int2(a, b));
  int cls_begin[10];
  sort_pairs_by_class(sh, &pr, cls_begin);
  const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
  DevBuf dsh, dex, dcc, dpr, dax, daz, dS, dT, dV;
  HIPC(dsh.ensure(sizeof(ShellDev) * ns));
  HIPC(dex.ensure(sizeof(double) * ctx->exps.size()));
  HIPC(dcc.ensure(sizeof(double) * ctx->ccoef.size()));
  HIPC(dpr.ensure(sizeof(int2) * pr.size()));
  HIPC(dax.ensure(sizeof(double) * 3 * std::max(natom, 1)));
  HIPC(daz.ensure(sizeof(double) * std::max(natom, 1)));
  HIPC(dS.ensure(n2 * sizeof(double)));
  HIPC(dT.ensure(n2 * sizeof(double)));
  HIPC(dV.ensure(n2 * sizeof(double)));
  HIPC(hipMemcpyAsync(dsh.p, sh.data(), sizeof(ShellDev) * ns, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dex.p, ctx->exps.data(), sizeof(double) * ctx->exps.size(),
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dcc.p, ctx->ccoef.data(), sizeof(double) * ctx->ccoef.size(),
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dpr.p, pr.data(), sizeof(int2) * pr.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  if (natom > 0) {
    HIPC(hipMemcpyAsync(dax.p, xyz, sizeof(double) * 3 * natom, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(daz.p, Z, sizeof(double) * natom, hipMemcpyHostToDevice, ctx->stream));
  }
  OneElArgs a{};
  a.shells = dsh.as<ShellDev>();
  a.exps = dex.as<double>();
  a.ccoef = dcc.as<double>();
  a.pairs = dpr.as<int2>();
  a.npairs = static_cast<int>(pr.size());
  a.nbf = ctx->nbf;
  a.natom = natom;
  a.atom_xyz = dax.as<double>();
  a.atom_z = daz.as<double>();
  a.boys_table = ctx->boys_d.as<double>();
  a.S = dS.as<double>();
  a.T = dT.as<double>();
  a.V = dV.as<double>();
  for (int c = 0; c < 10; ++c) a.cls_begin[c] = cls_begin[c];
  HIPC(one_electron_launch(a, ctx->stream));
  if (S) HIPC(hipMemcpyAsync(S, dS.p, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (T) HIPC(hipMemcpyAsync(T, dT.p, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (V) HIPC(hipMemcpyAsync(V, dV.p, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return RCHEM_OK;
}

int rchem_scf_opts_default(rchem_scf_opts* o) {
  if (!o) return RCHEM_ERR_INVALID;
  o->nocc = 0;
  o->max_iter = 1024;
  o->thresh_e = 1e-11;
  rchem_opts_default(&o->jk);
  o->verbose = 0;
  o->diis = 0;
  o->graph = 0;
  o->purify = 0;
  o->damping = 0.0;
  o->damping_until = 0.0;
  return RCHEM_OK;
}

}  // extern "C"

namespace {

// An SCF iteration's Fock build -- jk_device (planner, class kernels,
// digestion) and F = H + 2J - K -- captured once as a hipGraph and replayed.
// Every buffer it touches keeps its address between iterations (D is rewritten
// in place), the planner recomputes everything density-dependent on the
// device, and its fills are kernels, so a replay is the eager build.  The first
// build of a configuration runs eagerly (host planning, allocations); the next
// one is captured; a change of precision recaptures.  Per-class timing (event
// records around each launch) keeps the build eager.
struct FockGraph {
  bool enabled = false;
  int eager_runs = 0;
  int precision = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~FockGraph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    eager_runs = 0;
  }
  int run(rchem_ctx* ctx, const rchem_opts& jk, const double* D, double* J, double* K,
          const double* H, double* F, size_t n2) {
    auto eager = [&]() -> int {
      TRY(jk_device(ctx, D, J, K, &jk));
      if (ctx->nmol > 1 && ctx->mol_nmax > 0)  // a multi-molecule basis: its diagonal blocks
        HIPC(batch_fock(H, J, K, F, static_cast<size_t>(ctx->nbf), ctx->mol_f0_d.as<int>(),
                        ctx->mol_nb_d.as<int>(), ctx->nmol, ctx->mol_nmax, ctx->stream));
      else
        HIPC(fock_launch(H, J, K, F, n2, ctx->stream));
      return RCHEM_OK;
    };
    if (!enabled || ctx->profiling) return eager();
    if (jk.precision != precision) {
      reset();
      precision = jk.precision;
    }
    if (exec) {
      HIPC(hipGraphLaunch(exec, ctx->stream));
      return RCHEM_OK;
    }
    if (eager_runs == 0) {
      ++eager_runs;
      return eager();
    }
    HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    const int rc = eager();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    if (rc != RCHEM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPC(ec);
    graph = g;
    HIPC(hipGraphInstantiateWithFlags(&exec, graph, 0));
    HIPC(hipGraphLaunch(exec, ctx->stream));
    return RCHEM_OK;
  }
};

int scf_run(rchem_ctx* ctx, int natom, const double* Z, const double* xyz,
            const rchem_scf_opts* opts, rchem_scf_result* res, double dconv) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!ctx->have_basis) return ctx->fail(RCHEM_ERR_NO_BASIS, "no basis uploaded");
  if (!opts || !res) return ctx->fail(RCHEM_ERR_INVALID, "null opts or result");
  const int n = ctx->nbf;
  if (opts->nocc <= 0 || opts->nocc > n)
    return ctx->fail(RCHEM_ERR_INVALID, "nocc out of range");
  TRY(check_opts(ctx, &opts->jk));
  TRY(check_scf_opts(ctx, opts));
  TRY(set_device(ctx));
  const size_t n2 = static_cast<size_t>(n) * n;
  // one-electron matrices on the host path of this library, then resident on device
  std::vector<double> S(n2), T(n2), V(n2), H(n2);
  TRY(rchem_one_electron(ctx, natom, Z, xyz, S.data(), T.data(), V.data()));
  for (size_t i = 0; i < n2; ++i) H[i] = T[i] + V[i];
  DevBuf dS, dH, dX, dW, dFp, dJ, dK, dtmp, dC, dw, de, dinfo, dE;
  DevBuf& dF = ctx->scf_F;
  DevBuf& dD = ctx->scf_D;
  DevBuf& dDo = ctx->scf_Dold;
  for (DevBuf* b : {&dS, &dH, &dX, &dW, &dF, &dFp, &dC, &dD, &dJ, &dK, &dtmp, &dDo})
    HIPC(b->ensure(n2 * sizeof(double)));
  HIPC(dw.ensure(n * sizeof(double)));
  HIPC(de.ensure(n * sizeof(double)));
  HIPC(dinfo.ensure(sizeof(rocblas_int)));
  HIPC(dE.ensure(sizeof(double)));
  HIPC(hipMemcpyAsync(dS.p, S.data(), n2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dH.p, H.data(), n2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  const double one = 1.0, zero = 0.0;
  rocblas_handle hb = ctx->blas;
  // Loewdin X = U lambda^-1/2 U^T (rchem.rs:44-51); symmetric, so the
  // row/column-major distinction does not matter for any matrix below.
  HIPC(hipMemcpyAsync(dtmp.p, dS.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  BLASC(rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, dtmp.as<double>(), n,
                         dw.as<double>(), de.as<double>(), dinfo.as<rocblas_int>()));
  HIPC(scale_columns_launch(dtmp.as<double>(), dw.as<double>(), dW.as<double>(), n, ctx->stream));
  BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one,
                      dW.as<double>(), n, dtmp.as<double>(), n, &zero, dX.as<double>(), n));
  // Canonical purification of F' (dFp) into P (dP): GEMMs until tr(P - P^2)
  // falls below kPurifyTol, then two more steps (each squares the distance of
  // every eigenvalue from 0 or 1); *steps = 0 if that does not happen in
  // kPurifyMax steps (no gap at the Fermi level), and the eigensolver takes over.
  constexpr int kPurifyMax = 200;
  constexpr double kPurifyTol = 1e-9;
  DevBuf dP, dP2, dP3, dpl, dph, dpar;
  if (opts->purify) {
    for (DevBuf* b : {&dP, &dP2, &dP3}) HIPC(b->ensure(n2 * sizeof(double)));
    HIPC(dpl.ensure(n * sizeof(double)));
    HIPC(dph.ensure(n * sizeof(double)));
    HIPC(dpar.ensure((4 + 2 * kPurifyMax) * sizeof(double)));
  }
  auto purify_step = [&](double* coef) -> int {
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                        dP.as<double>(), n, dP.as<double>(), n, &zero, dP2.as<double>(), n));
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                        dP2.as<double>(), n, dP.as<double>(), n, &zero, dP3.as<double>(), n));
    HIPC(purify_step_launch(dP.as<double>(), dP2.as<double>(), dP3.as<double>(), n, coef,
                            ctx->stream));
    return RCHEM_OK;
  };
  auto purify = [&](int* steps) -> int {
    *steps = 0;
    HIPC(purify_start_launch(dFp.as<double>(), n, opts->nocc, dpl.as<double>(), dph.as<double>(),
                             dpar.as<double>(), dP.as<double>(), ctx->stream));
    int extra = -1;
    for (int k = 0; k < kPurifyMax; ++k) {
      TRY(purify_step(dpar.as<double>() + 4));
      double t1 = 0.0;
      HIPC(hipMemcpyAsync(&t1, dpar.as<double>() + 5, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
      HIPC(hipStreamSynchronize(ctx->stream));
      // tr(P - P^2) = sum lambda (1 - lambda) >= 0 while the spectrum stays in [0, 1]
      if (!std::isfinite(t1) || t1 < -kPurifyTol) return RCHEM_OK;
      if (extra < 0 && t1 < kPurifyTol) extra = 2;
      if (extra >= 0 && extra-- == 0) {
        *steps = k + 1;
        return RCHEM_OK;
      }
    }
    return RCHEM_OK;
  };
  // F' = X^T Fm X (rchem.rs:56, 74)
  auto orthogonalise = [&](const double* Fm) -> int {
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, Fm, n,
                        dX.as<double>(), n, &zero, dtmp.as<double>(), n));
    BLASC(rocblas_dgemm(hb, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                        dX.as<double>(), n, dtmp.as<double>(), n, &zero, dFp.as<double>(), n));
    return RCHEM_OK;
  };
  // Dout = X P X^T from the purified P in dP
  auto density_from_p = [&](double* Dout) -> int {
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                        dX.as<double>(), n, dP.as<double>(), n, &zero, dtmp.as<double>(), n));
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one,
                        dtmp.as<double>(), n, dX.as<double>(), n, &zero, Dout, n));
    return RCHEM_OK;
  };
  // Dout = the density of Fm: F' = X^T F X, eigh, C = X C', D = C_occ C_occ^T
  // (rchem.rs:56-59, 74-78, 101-104), or X P X^T by purification; *psteps
  // (may be null) receives the purification's step count, 0 if the
  // eigensolver formed it
  auto density = [&](const double* Fm, double* Dout, int* psteps) -> int {
    if (psteps) *psteps = 0;
    TRY(orthogonalise(Fm));
    if (opts->purify) {
      int steps = 0;
      TRY(purify(&steps));
      if (steps > 0) {
        if (psteps) *psteps = steps;
        return density_from_p(Dout);
      }
    }
    BLASC(rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, dFp.as<double>(), n,
                           dw.as<double>(), de.as<double>(), dinfo.as<rocblas_int>()));
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                        dX.as<double>(), n, dFp.as<double>(), n, &zero, dC.as<double>(), n));
    BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_transpose, n, n,
                        opts->nocc, &one, dC.as<double>(), n, dC.as<double>(), n, &zero, Dout,
                        n));
    return RCHEM_OK;
  };
  auto energy = [&](const double* Fm, double* e) -> int {
    HIPC(energy_launch(dD.as<double>(), dH.as<double>(), Fm, dE.as<double>(), n2, ctx->stream));
    HIPC(hipMemcpyAsync(e, dE.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return RCHEM_OK;
  };
  // D <- (1 - a) D + a D_old (opts->damping), D_old saved by save_old()
  auto save_old = [&]() -> int {
    HIPC(hipMemcpyAsync(dDo.p, dD.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return RCHEM_OK;
  };
  auto damp = [&](double measure) -> int {
    if (opts->damping <= 0.0 || measure < opts->damping_until) return RCHEM_OK;
    const double keep = 1.0 - opts->damping, mix = opts->damping;
    const rocblas_int nn2 = static_cast<rocblas_int>(n2);
    BLASC(rocblas_dscal(hb, nn2, &keep, dD.as<double>(), 1));
    BLASC(rocblas_daxpy(hb, nn2, &mix, dDo.as<double>(), 1, dD.as<double>(), 1));
    return RCHEM_OK;
  };
  int seen_steps = 0;  // purification steps the last eager density needed
  TRY(density(dH.as<double>(), dD.as<double>(), &seen_steps));
  double e_new = 0.0;
  TRY(energy(dH.as<double>(), &e_new));
  res->converged = 0;
  res->iterations = 0;
  res->graph_iterations = 0;
  int it = 0;
  // Mixed precision: FP32 quartets until the energy changes by less than
  // kSwitch per iteration, FP64 for every quartet from then on, so the last
  // iterations (and the converged energy) see the all-FP64 Fock matrix.
  constexpr double kSwitch = 1e-7;
  rchem_opts jk = opts->jk;
  FockGraph fock_graph;
  res->fp64_from = jk.precision == RCHEM_PREC_MIXED ? -1 : 0;
  if (opts->diis > 0) {
    // Pulay DIIS on the device: extrapolate F from the last k Fock matrices so
    // that the orthogonalised commutators X^T (FDS - SDF) X combine to the
    // smallest norm.  With purified densities an iteration has no host step and
    // is replayed from one captured graph.
    const int k = std::min(opts->diis, kDiisMax);
    const double errtol = dconv > 0.0 ? dconv : 1e-8;
    DevBuf dFs, dEs, dEc, dFx, dDn, dpart, dst;
    HIPC(dFs.ensure(n2 * sizeof(double) * k));
    HIPC(dEs.ensure(n2 * sizeof(double) * k));
    for (DevBuf* b : {&dEc, &dFx, &dDn}) HIPC(b->ensure(n2 * sizeof(double)));
    HIPC(dpart.ensure(sizeof(double) * kDiisBlocks * (kDiisMax + 1)));
    HIPC(dst.ensure(sizeof(ScfDevState)));
    if (!ctx->scf_status_h) HIPC(hipHostMalloc(&ctx->scf_status_h, sizeof(ScfDevState)));
    ScfDevState* hst = ctx->scf_status_h;
    ScfDevState* st = dst.as<ScfDevState>();
    HIPC(hipMemsetAsync(dst.p, 0, sizeof(ScfDevState), ctx->stream));
    auto params = [&]() {
      ScfDevParams p{};
      p.thresh_e = opts->thresh_e;
      p.errtol = errtol;
      p.switch_de = kSwitch;
      p.damping = opts->damping;
      p.damping_until = opts->damping_until;
      p.k = k;
      p.mixed = jk.precision == RCHEM_PREC_MIXED;
      p.fp64 = jk.precision == RCHEM_PREC_FP64;
      return p;
    };
    // energy of the current D, commutator, DIIS push and solve, F_x
    auto front = [&]() -> int {
      HIPC(energy_launch(dD.as<double>(), dH.as<double>(), dF.as<double>(), dE.as<double>(), n2,
                         ctx->stream));
      // FDS - SDF, then X^T (.) X
      BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                          dF.as<double>(), n, dD.as<double>(), n, &zero, dtmp.as<double>(), n));
      BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                          dtmp.as<double>(), n, dS.as<double>(), n, &zero, dFp.as<double>(), n));
      HIPC(antisym_launch(dFp.as<double>(), dFx.as<double>(), n, ctx->stream));
      BLASC(rocblas_dgemm(hb, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                          dX.as<double>(), n, dFx.as<double>(), n, &zero, dtmp.as<double>(), n));
      BLASC(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                          dtmp.as<double>(), n, dX.as<double>(), n, &zero, dEc.as<double>(), n));
      HIPC(diis_push_launch(dEc.as<double>(), dF.as<double>(), dEs.as<double>(),
                            dFs.as<double>(), n2, st, k, dpart.as<double>(), ctx->stream));
      HIPC(diis_solve_launch(st, dpart.as<double>(), dE.as<double>(), params(), ctx->stream));
      HIPC(diis_extrapolate_launch(dFs.as<double>(), n2, st, dFx.as<double>(), ctx->stream));
      return RCHEM_OK;
    };
    // the density of F_x by a fixed number of purification steps, committed
    // by the device if tr(P - P^2) was below kPurifyTol two steps before the end
    auto back_fixed = [&](int steps) -> int {
      TRY(orthogonalise(dFx.as<double>()));
      HIPC(purify_start_launch(dFp.as<double>(), n, opts->nocc, dpl.as<double>(),
                               dph.as<double>(), dpar.as<double>(), dP.as<double>(), ctx->stream));
      double* coef = dpar.as<double>() + 4;
      for (int s = 0; s < steps; ++s) TRY(purify_step(coef + 2 * s));
      TRY(density_from_p(dDn.as<double>()));
      HIPC(scf_commit_launch(dD.as<double>(), dDn.as<double>(), n2, st, coef, steps, kPurifyTol,
                             opts->damping, ctx->stream));
      return RCHEM_OK;
    };
    // the density of F_x on the host's schedule (adaptive purification or the
    // eigensolver), committed unconditionally
    auto back_eager = [&](int* steps) -> int {
      TRY(density(dFx.as<double>(), dDn.as<double>(), steps));
      HIPC(scf_commit_launch(dD.as<double>(), dDn.as<double>(), n2, st, nullptr, 0, kPurifyTol,
                             opts->damping, ctx->stream));
      return RCHEM_OK;
    };
    auto read_status = [&]() -> int {
      HIPC(hipMemcpyAsync(hst, st, offsetof(ScfDevState, e_prev), hipMemcpyDeviceToHost,
                          ctx->stream));
      HIPC(hipStreamSynchronize(ctx->stream));
      return RCHEM_OK;
    };
    // whole-iteration graph: purified densities, graphs on, not profiling
    bool whole = opts->graph != 0 && opts->purify != 0 && !ctx->profiling && seen_steps > 0;
    fock_graph.enabled = opts->graph != 0 && !whole;
    hipGraph_t ig = nullptr;
    hipGraphExec_t ix = nullptr;
    // The graph holds a fixed number of purification steps: what the last
    // eager density or replay needed plus kStepMargin; a replay that needs more
    // raises purify_failed (its density is then formed eagerly and the graph
    // recaptured longer), one that needs kStepSlack fewer is recaptured shorter.
    constexpr int kStepMargin = 3, kStepSlack = 4;
    int ig_steps = 0;
    bool planned = false;  // a build ran at this precision (host planning done)
    auto drop_graph = [&](bool replan) {
      if (ix) (void)hipGraphExecDestroy(ix);
      if (ig) (void)hipGraphDestroy(ig);
      ix = nullptr;
      ig = nullptr;
      if (replan) planned = false;
    };
    auto iteration_body = [&](int steps) -> int {
      TRY(jk_device(ctx, dD.as<double>(), dJ.as<double>(), dK.as<double>(), &jk));
      HIPC(fock_launch(dH.as<double>(), dJ.as<double>(), dK.as<double>(), dF.as<double>(), n2,
                       ctx->stream));
      TRY(front());
      return back_fixed(steps);
    };
    int rc_loop = RCHEM_OK;
    for (; it < opts->max_iter; ++it) {
      bool replayed = false;
      if (whole && !ix && planned) {
        // capture (nothing runs), then replay; a refused capture falls back to
        // launching kernel by kernel for the rest of the SCF
        ig_steps = std::min(std::max(seen_steps + kStepMargin, 8), kPurifyMax);
        HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        ctx->capturing = true;
        const int rc = iteration_body(ig_steps);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        ctx->capturing = false;
        hipError_t ei = hipErrorUnknown;
        if (rc == RCHEM_OK && ec == hipSuccess && g) ei = hipGraphInstantiateWithFlags(&ix, g, 0);
        if (ei == hipSuccess) {
          ig = g;
        } else {
          if (g) (void)hipGraphDestroy(g);
          ix = nullptr;
          (void)hipGetLastError();
          whole = false;
          fock_graph.enabled = opts->graph != 0;
        }
      }
      if (whole && ix) {
        HIPC(hipGraphLaunch(ix, ctx->stream));
        replayed = true;
        ++res->graph_iterations;
      } else {
        TRY(fock_graph.run(ctx, jk, dD.as<double>(), dJ.as<double>(), dK.as<double>(),
                           dH.as<double>(), dF.as<double>(), n2));
        planned = true;
        TRY(front());
        TRY(read_status());
        if (!hst->switched) {
          int steps = 0;
          TRY(back_eager(&steps));
          if (steps > 0) seen_steps = steps;
          if (opts->purify && steps == 0 && whole) {
            whole = false;  // no gap: the eigensolver's turn, the Fock build alone replayed
            fock_graph.enabled = opts->graph != 0;
          }
        }
      }
      TRY(read_status());
      if (replayed && hst->purify_failed) {
        // more steps than the graph holds (or no gap): this iteration's
        // density eagerly, then a recapture with the steps it needed
        int steps = 0;
        TRY(back_eager(&steps));
        TRY(read_status());
        drop_graph(false);
        if (steps > 0) {
          seen_steps = steps;
        } else {
          whole = false;
          fock_graph.enabled = opts->graph != 0;
        }
      } else if (replayed && hst->purify_steps + kStepMargin + kStepSlack <= ig_steps) {
        seen_steps = hst->purify_steps;
        drop_graph(false);
      }
      const double e_cur = hst->e_cur;
      if (!std::isfinite(e_cur)) {
        rc_loop = ctx->fail(RCHEM_ERR_NONFINITE, "non-finite SCF energy");
        break;
      }
      if (opts->verbose) std::printf("%4d %20.12f %20.12f %12.3e\n", it, e_cur, hst->de, hst->errmax);
      e_new = e_cur;
      if (hst->switched) {
        // the next Fock matrix is all-FP64 (D kept, the DIIS history too)
        jk.precision = RCHEM_PREC_FP64;
        res->fp64_from = it + 1;
        drop_graph(true);
        continue;
      }
      if (hst->converged) {
        // D is already the density of this iteration's F (no extrapolation)
        if (opts->verbose) std::printf("Convergence achieved!\n");
        res->converged = 1;
        ++it;
        break;
      }
    }
    drop_graph(true);
    if (rc_loop != RCHEM_OK) return rc_loop;
  }
  double e_prev_plain = e_new - 1.0;  // |dE| of the previous iteration (damping measure)
  fock_graph.enabled = opts->graph != 0;
  for (; opts->diis <= 0 && it < opts->max_iter; ++it) {
    TRY(fock_graph.run(ctx, jk, dD.as<double>(), dJ.as<double>(), dK.as<double>(),
                       dH.as<double>(), dF.as<double>(), n2));
    if (dconv > 0.0 || opts->damping > 0.0) TRY(save_old());
    TRY(density(dF.as<double>(), dD.as<double>(), nullptr));
    TRY(damp(std::fabs(e_new - e_prev_plain)));
    const double e_old = e_new;
    TRY(energy(dF.as<double>(), &e_new));
    e_prev_plain = e_old;
    if (!std::isfinite(e_new)) return ctx->fail(RCHEM_ERR_NONFINITE, "non-finite SCF energy");
    const double de_ = e_new - e_old;
    if (opts->verbose) std::printf("%4d %20.12f %20.12f\n", it, e_new, de_);
    double ddmax = 0.0;
    if (dconv > 0.0) {
      // max |D_new - D_old| (the gradient wants a self-consistent density)
      const double m1 = -1.0;
      rocblas_int imax = 0;
      BLASC(rocblas_daxpy(hb, static_cast<rocblas_int>(n2), &m1, dD.as<double>(), 1,
                          dDo.as<double>(), 1));
      BLASC(rocblas_idamax(hb, static_cast<rocblas_int>(n2), dDo.as<double>(), 1, &imax));
      HIPC(hipMemcpy(&ddmax, dDo.as<double>() + (imax - 1), sizeof(double),
                     hipMemcpyDeviceToHost));
      ddmax = std::fabs(ddmax);
    }
    if (jk.precision == RCHEM_PREC_MIXED && std::fabs(de_) < kSwitch) {
      jk.precision = RCHEM_PREC_FP64;
      res->fp64_from = it + 1;
      continue;
    }
    if (std::fabs(de_) < opts->thresh_e && ddmax <= dconv) {
      if (opts->verbose) std::printf("Convergence achieved!\n");
      res->converged = 1;
      ++it;
      break;
    }
  }
  res->iterations = it;
  res->e_elec = e_new;
  double enuc = 0.0;
  for (int a = 0; a < natom; ++a)
    for (int b = 0; b < a; ++b) {
      const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
                   dz = xyz[3 * a + 2] - xyz[3 * b + 2];
      enuc += Z[a] * Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  res->e_nuc = enuc;
  res->e_total = e_new + enuc;
  HIPC(hipStreamSynchronize(ctx->stream));
  if (!res->converged)
    return ctx->fail(RCHEM_ERR_NOT_CONVERGED, "SCF did not converge in max_iter iterations");
  return RCHEM_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// Basis functions per chunk of molecules (default; RCHEM_BATCH_NBF sets it, up
// to kBatchMaxNbfLimit, where n^2 doubles per matrix stay below 2^31).  A chunk
// is one multi-molecule basis with dense n x n matrices (~7 of them alive while
// it iterates, 3 more during its set-up): 24576 functions keep that under
// 48 GB, while larger chunks put more molecules into every launch (the
// 512-molecule STO-3G batch: 8192 -> 9.5 s, 16384 -> 8.6 s, one chunk -> 8.1 s).
constexpr int kBatchMaxNbf = 24576;
constexpr int kBatchMaxNbfLimit = 40000;
constexpr double kBatchPadF = 1.0e4;  // Fock diagonal of a padded block's unused rows

// One chunk of the batched SCF: molecules [m0, m1) of the caller's lists as one
// multi-molecule basis.  zoff/aoff: offsets of molecule m0 in Z and in atoms.
// Molecules [m0, m1) of the arrays given (their atoms from aoff on) as one
// multi-molecule basis; mol_id[m] is molecule m's index in the caller's input
// (for messages).
int batch_chunk(rchem_ctx* ctx, int m0, int m1, const int* natom, const int* Z,
                const double* xyz, int aoff, const std::string& want, const rchem_scf_opts* opts,
                rchem_scf_result* res, const int* mol_id) {
  const int B = m1 - m0;
  auto lower = [](std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
  };
  ctx->shells.clear();
  ctx->exps.clear();
  ctx->coefs.clear();
  ctx->shell_mol.clear();
  std::vector<int2> mol_atoms(B);
  std::vector<int> nocc(B), mfirst_shell(B);
  std::vector<double> Zc, xyzc;
  int a = aoff;
  for (int b = 0; b < B; ++b) {
    mol_atoms[b] = int2{static_cast<int>(Zc.size()), 0};
    mfirst_shell[b] = static_cast<int>(ctx->shells.size());
    int ne = 0;
    for (int k = 0; k < natom[m0 + b]; ++k, ++a) {
      bool found = false;
      for (int r = 0; r < kBasisNumRows; ++r) {
        const BasisShellRow& row = kBasisRows[r];
        if (lower(row.set) != want || row.Z != Z[a]) continue;
        found = true;
        for (int q = 0; q < 2; ++q) {
          const int l = row.am[q];
          if (l < 0) continue;
          Shell sh;
          sh.l = l;
          sh.nprim = row.nprim;
          sh.p0 = static_cast<int>(ctx->exps.size());
          for (int x = 0; x < 3; ++x) sh.xyz[x] = xyz[3 * a + x];
          for (int p = 0; p < row.nprim; ++p) {
            ctx->exps.push_back(row.exps[p]);
            ctx->coefs.push_back(row.coefs[q][p]);
          }
          ctx->shells.push_back(sh);
          ctx->shell_mol.push_back(b);
        }
      }
      if (!found)
        return ctx->fail(RCHEM_ERR_INVALID, "element Z=" + std::to_string(Z[a]) + " not in basis");
      ne += Z[a];
      Zc.push_back(Z[a]);
      for (int x = 0; x < 3; ++x) xyzc.push_back(xyz[3 * a + x]);
    }
    mol_atoms[b].y = static_cast<int>(Zc.size());
    if (ne % 2 != 0)
      return ctx->fail(RCHEM_ERR_INVALID, "molecule " + std::to_string(mol_id[m0 + b]) +
                                              " has an odd electron count");
    nocc[b] = ne / 2;
  }
  ctx->nmol = B;
  TRY(finish_basis(ctx));
  const int n = ctx->nbf, ns = static_cast<int>(ctx->shells.size());
  const size_t n2 = static_cast<size_t>(n) * n;
  std::vector<int> f0(B), nb(B);
  int nmax = 1;
  for (int b = 0; b < B; ++b) {
    f0[b] = ctx->shells[mfirst_shell[b]].f0;
    const int end = b + 1 < B ? ctx->shells[mfirst_shell[b + 1]].f0 : n;
    nb[b] = end - f0[b];
    nmax = std::max(nmax, nb[b]);
    if (nocc[b] > nb[b])
      return ctx->fail(RCHEM_ERR_INVALID, "molecule with more electron pairs than functions");
  }
  hipStream_t st = ctx->stream;
  rocblas_handle hb = ctx->blas;
  // one-electron matrices of every molecule (pairs within a molecule, own nuclei)
  DevBuf dS, dT, dV, dH, dF, dD, dJ, dK;
  for (DevBuf* x : {&dS, &dT, &dV, &dH, &dF, &dD, &dJ, &dK}) {
    HIPC(x->ensure(n2 * sizeof(double)));
    HIPC(fill_launch(x->as<double>(), n2, 0.0, st));
  }
  {
    std::vector<ShellDev> sh(ns);
    for (int s2 = 0; s2 < ns; ++s2) {
      for (int x = 0; x < 3; ++x) sh[s2].xyz[x] = ctx->shells[s2].xyz[x];
      sh[s2].l = ctx->shells[s2].l;
      sh[s2].nprim = ctx->shells[s2].nprim;
      sh[s2].p0 = ctx->shells[s2].p0;
      sh[s2].f0 = ctx->shells[s2].f0;
    }
    std::vector<int2> pr;
    for (int x = 0; x < ns; ++x)
      for (int y = 0; y <= x; ++y)
        if (ctx->shell_mol[x] == ctx->shell_mol[y]) pr.push_back(make_int2(x, y));
    int cls_begin[10];
    sort_pairs_by_class(sh, &pr, cls_begin);
    DevBuf dsh, dex, dcc, dpr, dax, daz, dsm, dma;
    HIPC(dsh.ensure(sizeof(ShellDev) * ns));
    HIPC(dex.ensure(sizeof(double) * ctx->exps.size()));
    HIPC(dcc.ensure(sizeof(double) * ctx->ccoef.size()));
    HIPC(dpr.ensure(sizeof(int2) * pr.size()));
    HIPC(dax.ensure(sizeof(double) * xyzc.size()));
    HIPC(daz.ensure(sizeof(double) * Zc.size()));
    HIPC(dsm.ensure(sizeof(int) * ns));
    HIPC(dma.ensure(sizeof(int2) * B));
    HIPC(hipMemcpyAsync(dsh.p, sh.data(), sizeof(ShellDev) * ns, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dex.p, ctx->exps.data(), sizeof(double) * ctx->exps.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dcc.p, ctx->ccoef.data(), sizeof(double) * ctx->ccoef.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dpr.p, pr.data(), sizeof(int2) * pr.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dax.p, xyzc.data(), sizeof(double) * xyzc.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(daz.p, Zc.data(), sizeof(double) * Zc.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dsm.p, ctx->shell_mol.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dma.p, mol_atoms.data(), sizeof(int2) * B, hipMemcpyHostToDevice, st));
    OneElArgs oa{};
    oa.shells = dsh.as<ShellDev>();
    oa.exps = dex.as<double>();
    oa.ccoef = dcc.as<double>();
    oa.pairs = dpr.as<int2>();
    oa.npairs = static_cast<int>(pr.size());
    oa.nbf = n;
    oa.natom = static_cast<int>(Zc.size());
    oa.atom_xyz = dax.as<double>();
    oa.atom_z = daz.as<double>();
    oa.boys_table = ctx->boys_d.as<double>();
    oa.S = dS.as<double>();
    oa.T = dT.as<double>();
    oa.V = dV.as<double>();
    oa.shell_mol = dsm.as<int>();
    oa.mol_atoms = dma.as<int2>();
    for (int c = 0; c < 10; ++c) oa.cls_begin[c] = cls_begin[c];
    HIPC(one_electron_launch(oa, st));
    HIPC(hipStreamSynchronize(st));
  }
  const double one = 1.0, zero = 0.0;
  const rocblas_int nn2 = static_cast<rocblas_int>(n2);
  HIPC(hipMemcpyAsync(dH.p, dT.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
  BLASC(rocblas_daxpy(hb, nn2, &one, dV.as<double>(), 1, dH.as<double>(), 1));
  HIPC(hipStreamSynchronize(st));
  dT.release();  // H = T + V is all the iterations need of them
  dV.release();
  // padded per-molecule batches.  The batched linear algebra runs over the
  // molecules still iterating ("slots", in molecule order): when molecules
  // retire, the slots' per-molecule state (S, X, the DIIS history) is
  // gathered into a smaller batch, padded to the largest remaining molecule,
  // so that the iterations left to the slowest molecules cost in proportion
  // to them, not to the whole chunk.
  int Bs = B, nm = nmax;
  std::vector<int> slot_mol(B);
  for (int b = 0; b < B; ++b) slot_mol[b] = b;
  size_t m2 = static_cast<size_t>(nmax) * nmax;
  rocblas_stride sm = static_cast<rocblas_stride>(m2);
  DevBuf df0, dnb, dnocc, dSp, dX, dW, dT1, dFb, dC, dDb, dw, de, dinfo, dEm, dslots, dgather;
  HIPC(df0.ensure(sizeof(int) * B));
  HIPC(dnb.ensure(sizeof(int) * B));
  // the Fock builds of this basis touch only the molecules' diagonal blocks
  HIPC(ctx->mol_f0_d.ensure(sizeof(int) * B));
  HIPC(ctx->mol_nb_d.ensure(sizeof(int) * B));
  HIPC(hipMemcpyAsync(ctx->mol_f0_d.p, f0.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(ctx->mol_nb_d.p, nb.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  ctx->mol_nmax = nmax;
  HIPC(dnocc.ensure(sizeof(int) * B));
  for (DevBuf* x : {&dSp, &dX, &dW, &dT1, &dFb, &dC, &dDb}) HIPC(x->ensure(sizeof(double) * m2 * B));
  HIPC(dw.ensure(sizeof(double) * nmax * B));
  HIPC(de.ensure(sizeof(double) * nmax * B));
  HIPC(dinfo.ensure(sizeof(rocblas_int) * B));
  HIPC(dEm.ensure(sizeof(double) * B));
  HIPC(dslots.ensure(sizeof(int) * B));
  HIPC(hipMemcpyAsync(df0.p, f0.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(dnb.p, nb.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(dnocc.p, nocc.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  auto gemm_b = [&](rocblas_operation ta, rocblas_operation tb, const double* A, const double* Bm,
                    double* C) {
    return rocblas_dgemm_strided_batched(hb, ta, tb, nm, nm, nm, &one, A, nm, sm, Bm, nm, sm,
                                         &zero, C, nm, sm, Bs);
  };
  // Loewdin X_b = U_b lambda_b^-1/2 U_b^T (rchem.rs:44-51), padding = identity
  HIPC(batch_pack(dS.as<double>(), n, df0.as<int>(), dnb.as<int>(), nm, Bs, 1.0, dSp.as<double>(), st));
  HIPC(hipMemcpyAsync(dT1.p, dSp.p, sizeof(double) * m2 * Bs, hipMemcpyDeviceToDevice, st));
  BLASC(rocsolver_dsyevd_strided_batched(hb, rocblas_evect_original, rocblas_fill_upper, nm,
                                         dT1.as<double>(), nm, sm, dw.as<double>(), nm,
                                         de.as<double>(), nm, dinfo.as<rocblas_int>(), Bs));
  HIPC(batch_inv_sqrt_cols(dT1.as<double>(), dw.as<double>(), nm, Bs, dW.as<double>(), st));
  BLASC(gemm_b(rocblas_operation_none, rocblas_operation_transpose, dW.as<double>(),
               dT1.as<double>(), dX.as<double>()));
  HIPC(hipStreamSynchronize(st));
  dS.release();  // its blocks live on in dSp
  // density from a padded Fock batch: F' = X^T F X, eigh, C = X C', D = C_occ C_occ^T
  auto density_b = [&](const double* Fb) -> int {
    BLASC(gemm_b(rocblas_operation_none, rocblas_operation_none, Fb, dX.as<double>(), dT1.as<double>()));
    BLASC(gemm_b(rocblas_operation_transpose, rocblas_operation_none, dX.as<double>(),
                 dT1.as<double>(), dW.as<double>()));
    BLASC(rocsolver_dsyevd_strided_batched(hb, rocblas_evect_original, rocblas_fill_upper, nm,
                                           dW.as<double>(), nm, sm, dw.as<double>(), nm,
                                           de.as<double>(), nm, dinfo.as<rocblas_int>(), Bs));
    BLASC(gemm_b(rocblas_operation_none, rocblas_operation_none, dX.as<double>(), dW.as<double>(),
                 dC.as<double>()));
    HIPC(batch_mask_cols(dC.as<double>(), nm, dnocc.as<int>(), Bs, st));
    BLASC(gemm_b(rocblas_operation_none, rocblas_operation_transpose, dC.as<double>(),
                 dC.as<double>(), dDb.as<double>()));
    HIPC(batch_unpack(dDb.as<double>(), nm, df0.as<int>(), dnb.as<int>(), Bs, n, dD.as<double>(), st));
    return RCHEM_OK;
  };
  auto density_dense = [&](const double* Fd) -> int {
    HIPC(batch_pack(Fd, n, df0.as<int>(), dnb.as<int>(), nm, Bs, kBatchPadF, dFb.as<double>(), st));
    return density_b(dFb.as<double>());
  };
  // per-slot energies into the molecules' entries of e
  std::vector<double> eslot(B);
  auto energies = [&](const double* Fd, std::vector<double>* e) -> int {
    HIPC(batch_energy(dD.as<double>(), dH.as<double>(), Fd, n, df0.as<int>(), dnb.as<int>(), Bs,
                      dEm.as<double>(), st));
    HIPC(hipMemcpyAsync(eslot.data(), dEm.p, sizeof(double) * Bs, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int s = 0; s < Bs; ++s) (*e)[slot_mol[s]] = eslot[s];
    return RCHEM_OK;
  };
  std::vector<double> e_cur(B, 0.0), e_prev(B, 0.0), errmax(B, 0.0), de_last(B, 1.0);
  TRY(density_dense(dH.as<double>()));  // core-Hamiltonian guess (rchem.rs:56-61)
  TRY(energies(dH.as<double>(), &e_cur));
  std::vector<int> conv_at(B, -1);
  std::vector<double> e_final(B, 0.0);
  // converged molecules retire: the planner skips their quartets from the
  // next Fock build on, and their energy is kept from the iteration that
  // converged (their later J, K, F and D are no longer meaningful)
  // per molecule: 1 active, 2 active past its own switch to all-FP64 quartets
  // (mixed precision; each molecule switches once its energy changes by less
  // than 1e-7 Eh), 0 retired
  std::vector<int> active(B, opts->jk.precision == RCHEM_PREC_MIXED ? 1 : 2);
  std::vector<int> fp64_at(B, opts->jk.precision == RCHEM_PREC_MIXED ? -1 : 0);
  HIPC(ctx->mol_active_d.ensure(sizeof(int) * B));
  HIPC(hipMemcpyAsync(ctx->mol_active_d.p, active.data(), sizeof(int) * B, hipMemcpyHostToDevice,
                      st));
  HIPC(hipStreamSynchronize(st));
  ctx->mol_active_on = true;
  struct ActiveOff {
    rchem_ctx* c;
    ~ActiveOff() { c->mol_active_on = false; }
  } active_off{ctx};
  rchem_opts jk = opts->jk;
  const int k = std::min(std::max(opts->diis, 0), 32);
  DevBuf dFs, dEs, dDp, dE1, dpairs, ddots, dcoef, derr;
  if (k > 0) {
    HIPC(dFs.ensure(sizeof(double) * m2 * B * k));
    HIPC(dEs.ensure(sizeof(double) * m2 * B * k));
    HIPC(dDp.ensure(sizeof(double) * m2 * B));
    HIPC(dE1.ensure(sizeof(double) * m2 * B));
    HIPC(dpairs.ensure(sizeof(long long) * 2 * B * k));
    HIPC(ddots.ensure(sizeof(double) * B * k));
    HIPC(dcoef.ensure(sizeof(double) * B * k));
    HIPC(derr.ensure(sizeof(double) * B));
  }
  // DIIS matrices per molecule (entries <e_i, e_j> of its ring)
  std::vector<double> Bm(static_cast<size_t>(B) * k * k, 0.0);
  // density damping (opts->damping): the slots' previous densities (padded)
  // and per-slot factors
  DevBuf dDold, dalpha;
  std::vector<double> alpha(B, 0.0), errslot(B, 0.0);
  const bool damping = opts->damping > 0.0;
  if (damping) {
    HIPC(dDold.ensure(sizeof(double) * m2 * B));
    HIPC(dalpha.ensure(sizeof(double) * B));
  }
  auto save_old = [&]() -> int {
    if (damping)
      HIPC(batch_pack(dD.as<double>(), n, df0.as<int>(), dnb.as<int>(), nm, Bs, 0.0,
                      dDold.as<double>(), st));
    return RCHEM_OK;
  };
  // measure: per molecule
  auto damp_b = [&](const std::vector<double>& measure) -> int {
    if (!damping) return RCHEM_OK;
    for (int s = 0; s < Bs; ++s) {
      const int b = slot_mol[s];
      alpha[s] = (conv_at[b] < 0 && measure[b] >= opts->damping_until) ? opts->damping : 0.0;
    }
    HIPC(hipMemcpyAsync(dalpha.p, alpha.data(), sizeof(double) * Bs, hipMemcpyHostToDevice, st));
    HIPC(batch_blend(dD.as<double>(), dDold.as<double>(), n, df0.as<int>(), dnb.as<int>(),
                     dalpha.as<double>(), nm, Bs, st));
    HIPC(hipStreamSynchronize(st));  // alpha is reused next iteration
    return RCHEM_OK;
  };
  // drops the retired molecules' slots: gathers S, X and the DIIS history of
  // the remaining ones, padded to the largest of them (the padding rows and
  // columns hold the identity, pad * identity or zeros, so the top-left part
  // of a block is the block at the smaller size)
  auto compact = [&]() -> int {
    std::vector<int> keep;
    int nm_new = 1;
    for (int s = 0; s < Bs; ++s)
      if (conv_at[slot_mol[s]] < 0) {
        keep.push_back(s);
        nm_new = std::max(nm_new, nb[slot_mol[s]]);
      }
    const int Bn = static_cast<int>(keep.size());
    if (Bn == Bs || Bn == 0) return RCHEM_OK;
    const size_t m2n = static_cast<size_t>(nm_new) * nm_new;
    HIPC(hipMemcpyAsync(dslots.p, keep.data(), sizeof(int) * Bn, hipMemcpyHostToDevice, st));
    auto gather = [&](DevBuf& x, int nring) -> int {
      HIPC(dgather.ensure(sizeof(double) * m2n * Bn * nring));
      HIPC(batch_gather(x.as<double>(), m2 * Bs, nm, dslots.as<int>(), Bn, nring,
                        dgather.as<double>(), m2n * Bn, nm_new, st));
      HIPC(hipMemcpyAsync(x.p, dgather.p, sizeof(double) * m2n * Bn * nring,
                          hipMemcpyDeviceToDevice, st));
      return RCHEM_OK;
    };
    TRY(gather(dSp, 1));
    TRY(gather(dX, 1));
    if (k > 0) {
      TRY(gather(dFs, k));
      TRY(gather(dEs, k));
    }
    std::vector<int> mols(Bn), sf0(Bn), snb(Bn), socc(Bn);
    for (int s = 0; s < Bn; ++s) {
      mols[s] = slot_mol[keep[s]];
      sf0[s] = f0[mols[s]];
      snb[s] = nb[mols[s]];
      socc[s] = nocc[mols[s]];
    }
    HIPC(hipMemcpyAsync(df0.p, sf0.data(), sizeof(int) * Bn, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dnb.p, snb.data(), sizeof(int) * Bn, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(dnocc.p, socc.data(), sizeof(int) * Bn, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));  // keep and the per-slot vectors are host temporaries
    slot_mol = mols;
    Bs = Bn;
    nm = nm_new;
    m2 = m2n;
    sm = static_cast<rocblas_stride>(m2);
    return RCHEM_OK;
  };
  int count = 0, it = 0;
  bool all_conv = false;
  FockGraph fock_graph;
  fock_graph.enabled = opts->graph != 0;
  for (; it < opts->max_iter; ++it) {
    TRY(fock_graph.run(ctx, jk, dD.as<double>(), dJ.as<double>(), dK.as<double>(),
                       dH.as<double>(), dF.as<double>(), n2));
    e_prev = e_cur;
    if (k > 0) {
      // energy of the current density with its own Fock matrix
      TRY(energies(dF.as<double>(), &e_cur));
      const int ring = count % k;
      double* Fslot = dFs.as<double>() + m2 * Bs * ring;
      double* Eslot = dEs.as<double>() + m2 * Bs * ring;
      HIPC(batch_pack(dF.as<double>(), n, df0.as<int>(), dnb.as<int>(), nm, Bs, kBatchPadF, Fslot, st));
      HIPC(batch_pack(dD.as<double>(), n, df0.as<int>(), dnb.as<int>(), nm, Bs, 0.0, dDp.as<double>(), st));
      BLASC(gemm_b(rocblas_operation_none, rocblas_operation_none, Fslot, dDp.as<double>(), dT1.as<double>()));
      BLASC(gemm_b(rocblas_operation_none, rocblas_operation_none, dT1.as<double>(), dSp.as<double>(), dW.as<double>()));
      HIPC(batch_antisym(dW.as<double>(), dE1.as<double>(), nm, Bs, st));
      BLASC(gemm_b(rocblas_operation_transpose, rocblas_operation_none, dX.as<double>(), dE1.as<double>(), dT1.as<double>()));
      BLASC(gemm_b(rocblas_operation_none, rocblas_operation_none, dT1.as<double>(), dX.as<double>(), Eslot));
      ++count;
      const int m = std::min(count, k);
      std::vector<long long> prs(static_cast<size_t>(2) * Bs * m);
      for (int s = 0; s < Bs; ++s)
        for (int i = 0; i < m; ++i) {
          prs[2 * (static_cast<size_t>(s) * m + i)] = static_cast<long long>(m2 * Bs * ring + m2 * s);
          prs[2 * (static_cast<size_t>(s) * m + i) + 1] = static_cast<long long>(m2 * Bs * i + m2 * s);
        }
      HIPC(hipMemcpyAsync(dpairs.p, prs.data(), sizeof(long long) * prs.size(), hipMemcpyHostToDevice, st));
      HIPC(batch_dots(dEs.as<double>(), dpairs.as<long long>(), Bs * m, m2, ddots.as<double>(), st));
      HIPC(batch_maxabs(Eslot, nm, Bs, derr.as<double>(), st));
      std::vector<double> dots(static_cast<size_t>(Bs) * m);
      HIPC(hipMemcpyAsync(dots.data(), ddots.p, sizeof(double) * dots.size(), hipMemcpyDeviceToHost, st));
      HIPC(hipMemcpyAsync(errslot.data(), derr.p, sizeof(double) * Bs, hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      for (int s = 0; s < Bs; ++s) {
        const int b = slot_mol[s];
        errmax[b] = errslot[s];
        double* Bb = Bm.data() + static_cast<size_t>(b) * k * k;
        for (int i = 0; i < m; ++i)
          Bb[ring * k + i] = Bb[i * k + ring] = dots[static_cast<size_t>(s) * m + i];
      }
    } else {
      TRY(save_old());
      TRY(density_dense(dF.as<double>()));
      TRY(damp_b(de_last));  // |dE| of the previous iteration
      TRY(energies(dF.as<double>(), &e_cur));  // rchem's sum D_new o (H + F(D_old))
      for (int s = 0; s < Bs; ++s) {
        const int b = slot_mol[s];
        de_last[b] = std::fabs(e_cur[b] - e_prev[b]);
      }
    }
    double demax = 0.0;
    bool changed = false;
    int mixed_left = 0;
    for (int b = 0; b < B; ++b) {
      if (conv_at[b] >= 0) continue;
      const double d = std::fabs(e_cur[b] - e_prev[b]);
      demax = std::max(demax, d);
      // per-molecule convergence-triggered switch to all-FP64 quartets
      // (every molecule by iteration 64, so a stalled one cannot stay in FP32)
      if (active[b] == 1 && it > 0 && (d < 1e-7 || it >= 64)) {
        active[b] = 2;
        fp64_at[b] = it + 1;
        changed = true;
      }
      if (active[b] == 1) ++mixed_left;
    }
    if (opts->verbose) std::printf("%4d batch max|dE| %12.3e (%d molecules iterating)\n", it, demax, Bs);
    if (jk.precision == RCHEM_PREC_MIXED && mixed_left == 0)
      jk.precision = RCHEM_PREC_FP64;  // no FP32 quartets left: drop the FP32 kernels
    all_conv = true;
    bool retired = false;
    for (int b = 0; b < B; ++b) {
      if (conv_at[b] >= 0) continue;
      // converged: its energy of this iteration came from an all-FP64 Fock matrix
      const bool ok = it > 0 && fp64_at[b] >= 0 && fp64_at[b] <= it &&
                      std::fabs(e_cur[b] - e_prev[b]) < opts->thresh_e &&
                      (k <= 0 || errmax[b] < 1e-8);
      if (ok) {
        conv_at[b] = it + 1;
        e_final[b] = e_cur[b];
        active[b] = 0;
        retired = true;
      } else {
        all_conv = false;
      }
    }
    if (all_conv) {
      ++it;
      break;
    }
    if (retired || changed)
      HIPC(hipMemcpyAsync(ctx->mol_active_d.p, active.data(), sizeof(int) * B,
                          hipMemcpyHostToDevice, st));
    if (retired) TRY(compact());
    if (k > 0) {
      const int m = std::min(count, k);
      std::vector<double> coef(static_cast<size_t>(Bs) * k, 0.0);
      for (int s = 0; s < Bs; ++s) {
        const double* Bb = Bm.data() + static_cast<size_t>(slot_mol[s]) * k * k;
        const int sz = m + 1;
        std::vector<double> A(static_cast<size_t>(sz) * sz, 0.0), rhs(sz, 0.0);
        for (int i = 0; i < m; ++i) {
          for (int j = 0; j < m; ++j) A[i * sz + j] = Bb[i * k + j];
          A[i * sz + m] = A[m * sz + i] = 1.0;
        }
        rhs[m] = 1.0;
        bool ok = true;
        for (int c = 0; c < sz && ok; ++c) {
          int piv = c;
          for (int r = c + 1; r < sz; ++r)
            if (std::fabs(A[r * sz + c]) > std::fabs(A[piv * sz + c])) piv = r;
          if (std::fabs(A[piv * sz + c]) < 1e-300) { ok = false; break; }
          if (piv != c) {
            for (int j = 0; j < sz; ++j) std::swap(A[c * sz + j], A[piv * sz + j]);
            std::swap(rhs[c], rhs[piv]);
          }
          for (int r = 0; r < sz; ++r) {
            if (r == c) continue;
            const double f = A[r * sz + c] / A[c * sz + c];
            for (int j = c; j < sz; ++j) A[r * sz + j] -= f * A[c * sz + j];
            rhs[r] -= f * rhs[c];
          }
        }
        const int last = (count - 1) % k;
        for (int i = 0; i < m; ++i)
          coef[static_cast<size_t>(s) * k + i] = ok ? rhs[i] / A[i * sz + i] : (i == last ? 1.0 : 0.0);
      }
      HIPC(hipMemcpyAsync(dcoef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, st));
      HIPC(batch_combine(dFs.as<double>(), m2 * Bs, m, k, dcoef.as<double>(), nm, Bs, dFb.as<double>(), st));
      TRY(save_old());
      TRY(density_b(dFb.as<double>()));
      TRY(damp_b(errmax));
    }
  }
  HIPC(hipStreamSynchronize(st));
  for (int b = 0; b < B; ++b) {
    rchem_scf_result& r = res[m0 + b];
    r.e_elec = conv_at[b] >= 0 ? e_final[b] : e_cur[b];
    double enuc = 0.0;
    for (int p = mol_atoms[b].x; p < mol_atoms[b].y; ++p)
      for (int q = mol_atoms[b].x; q < p; ++q) {
        double d2 = 0.0;
        for (int x = 0; x < 3; ++x) d2 += (xyzc[3 * p + x] - xyzc[3 * q + x]) * (xyzc[3 * p + x] - xyzc[3 * q + x]);
        enuc += Zc[p] * Zc[q] / std::sqrt(d2);
      }
    r.e_nuc = enuc;
    r.e_total = r.e_elec + enuc;
    r.iterations = conv_at[b] >= 0 ? conv_at[b] : it;
    r.converged = conv_at[b] >= 0 ? 1 : 0;
    r.fp64_from = fp64_at[b];
  }
  return all_conv ? RCHEM_OK : RCHEM_ERR_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int rchem_batch_scf(rchem_ctx* ctx, int nmol, const int* natom, const int* Z, const double* xyz,
                    const char* basis_name, const rchem_scf_opts* opts, rchem_scf_result* res) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (nmol <= 0 || !natom || !Z || !xyz || !basis_name || !opts || !res)
    return ctx->fail(RCHEM_ERR_INVALID, "rchem_batch_scf: bad arguments");
  TRY(check_opts(ctx, &opts->jk));
  TRY(check_scf_opts(ctx, opts));
  TRY(set_device(ctx));
  // the batch's own launch table while it runs (none: the context's)
  struct TableSwap {
    rchem_ctx* c;
    bool on;
    TableSwap(rchem_ctx* x) : c(x), on(!x->tuning_batch.empty()) {
      if (on) std::swap(c->tuning, c->tuning_batch);
    }
    ~TableSwap() {
      if (on) std::swap(c->tuning, c->tuning_batch);
      c->plan_tau = -1.0;
    }
  } table_swap(ctx);
  auto lower = [](std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
  };
  const std::string want = lower(basis_name);
  // functions per molecule, for chunking
  std::vector<int> nbf(nmol, 0);
  int a = 0;
  for (int m = 0; m < nmol; ++m) {
    if (natom[m] <= 0) return ctx->fail(RCHEM_ERR_INVALID, "molecule without atoms");
    for (int k = 0; k < natom[m]; ++k, ++a)
      for (int r = 0; r < kBasisNumRows; ++r)
        if (lower(kBasisRows[r].set) == want && kBasisRows[r].Z == Z[a])
          for (int q = 0; q < 2; ++q)
            if (kBasisRows[r].am[q] >= 0) nbf[m] += ncart(kBasisRows[r].am[q]);
  }
  // chunks of molecules of similar size (largest first): the padded batches
  // of a chunk are as large as its largest molecule, and molecules that take
  // many iterations (mostly the large ones) keep fewer small ones iterating
  // with them
  std::vector<int> order(nmol), astart(nmol);
  for (int m = 0, s = 0; m < nmol; ++m) {
    order[m] = m;
    astart[m] = s;
    s += natom[m];
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nbf[x] > nbf[y]; });
  std::vector<int> natom_s(nmol), Z_s;
  std::vector<double> xyz_s;
  Z_s.reserve(a);
  xyz_s.reserve(3 * static_cast<size_t>(a));
  for (int m = 0; m < nmol; ++m) {
    const int src = order[m];
    natom_s[m] = natom[src];
    for (int k = 0; k < natom[src]; ++k) {
      Z_s.push_back(Z[astart[src] + k]);
      for (int x = 0; x < 3; ++x) xyz_s.push_back(xyz[3 * static_cast<size_t>(astart[src] + k) + x]);
    }
  }
  std::vector<rchem_scf_result> res_s(nmol);
  int chunk_nbf = kBatchMaxNbf;
  if (const char* e = std::getenv("RCHEM_BATCH_NBF"))
    chunk_nbf = std::min(std::max(std::atoi(e), 1), kBatchMaxNbfLimit);
  // chunks of about equal size: the fewest chunks of at most chunk_nbf
  // functions, each filled to the mean
  long long total_nbf = 0;
  for (int m = 0; m < nmol; ++m) total_nbf += nbf[m];
  const long long nchunks = std::max<long long>(1, (total_nbf + chunk_nbf - 1) / chunk_nbf);
  const long long target = (total_nbf + nchunks - 1) / nchunks;
  int status = RCHEM_OK, m0 = 0, aoff = 0;
  while (m0 < nmol) {
    int m1 = m0, tot = 0, aend = aoff;
    while (m1 < nmol && (m1 == m0 || (tot < target && tot + nbf[order[m1]] <= chunk_nbf))) {
      tot += nbf[order[m1]];
      aend += natom_s[m1];
      ++m1;
    }
    const int rc = batch_chunk(ctx, m0, m1, natom_s.data(), Z_s.data(), xyz_s.data(), aoff, want,
                               opts, res_s.data(), order.data());
    if (rc != RCHEM_OK && rc != RCHEM_ERR_NOT_CONVERGED) {
      ctx->shell_mol.clear();
      ctx->nmol = 1;
      ctx->mol_nmax = 0;
      return rc;
    }
    if (rc == RCHEM_ERR_NOT_CONVERGED) status = rc;
    for (int m = m0; m < m1; ++m) res[order[m]] = res_s[m];
    m0 = m1;
    aoff = aend;
  }
  // the context's basis is the last chunk's; leave it as a plain basis
  ctx->shell_mol.clear();
  ctx->nmol = 1;
  ctx->mol_nmax = 0;
  ctx->have_basis = false;
  if (status == RCHEM_ERR_NOT_CONVERGED)
    return ctx->fail(RCHEM_ERR_NOT_CONVERGED, "some molecules did not converge in max_iter");
  return status;
}

int rchem_rhf_gradient(rchem_ctx* ctx, int natom, const double* Z, const double* xyz,
                       const rchem_scf_opts* opts, rchem_scf_result* res, double* grad) {
  if (!ctx) return RCHEM_ERR_INVALID;
  if (!grad) return ctx->fail(RCHEM_ERR_INVALID, "null gradient");
  if (natom <= 0 || !Z || !xyz) return ctx->fail(RCHEM_ERR_INVALID, "bad nuclei");
  const int rc = scf_run(ctx, natom, Z, xyz, opts, res, 1e-9);
  // The gradient formula holds at a stationary density only: without
  // convergence no gradient is formed and grad is left as the caller passed it.
  if (rc != RCHEM_OK) return rc;
  const int n = ctx->nbf, ns = static_cast<int>(ctx->shells.size());
  const size_t n2 = static_cast<size_t>(n) * n;
  // energy-weighted density W = C_occ diag(eps_occ) C_occ^T = D F D: the SCF
  // leaves D = C_occ C_occ^T, the density of F (F C = S C eps, C^T S C = 1,
  // so C_occ^T F C_occ = diag(eps_occ)); no orbitals needed
  DevBuf dDF, dW, dgrad, dpa, dsat, dsh, dex, dcc, dpr, dax, daz;
  HIPC(dDF.ensure(n2 * sizeof(double)));
  HIPC(dW.ensure(n2 * sizeof(double)));
  const double one = 1.0, zero = 0.0;
  BLASC(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                      ctx->scf_D.as<double>(), n, ctx->scf_F.as<double>(), n, &zero,
                      dDF.as<double>(), n));
  BLASC(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                      dDF.as<double>(), n, ctx->scf_D.as<double>(), n, &zero, dW.as<double>(),
                      n));
  // shell -> nucleus
  std::vector<int> satom(ns, -1);
  for (int s = 0; s < ns; ++s)
    for (int A = 0; A < natom; ++A) {
      const double dx = ctx->shells[s].xyz[0] - xyz[3 * A], dy = ctx->shells[s].xyz[1] - xyz[3 * A + 1],
                   dz = ctx->shells[s].xyz[2] - xyz[3 * A + 2];
      if (dx * dx + dy * dy + dz * dz < 1e-16) {
        satom[s] = A;
        break;
      }
    }
  for (int s = 0; s < ns; ++s)
    if (satom[s] < 0) return ctx->fail(RCHEM_ERR_INVALID, "a shell centre is not at a nucleus");
  HIPC(dgrad.ensure(sizeof(double) * 3 * natom));
  HIPC(fill_launch(dgrad.as<double>(), 3 * static_cast<size_t>(natom), 0.0, ctx->stream));
  // two-electron part: derivative quartets contracted with Gamma on the fly
  const double tau = std::min(opts->jk.schwarz_thresh > 0.0 ? opts->jk.schwarz_thresh : 1e-12,
                              1e-12) * 1e-2;
  TRY(prepare_host_tiles(ctx, tau));
  std::vector<int2> pat(ctx->pairs.size());
  for (size_t i = 0; i < pat.size(); ++i)
    pat[i] = make_int2(satom[ctx->pair_shells[i].first], satom[ctx->pair_shells[i].second]);
  HIPC(dpa.ensure(sizeof(int2) * std::max<size_t>(pat.size(), 1)));
  if (!pat.empty())
    HIPC(hipMemcpyAsync(dpa.p, pat.data(), sizeof(int2) * pat.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  DevBuf dgp;  // per-pair accumulators of the per-lane kernels
  const int npairs = static_cast<int>(ctx->pairs.size());
  HIPC(dgp.ensure(sizeof(double) * 6 * std::max(npairs, 1)));
  HIPC(fill_launch(dgp.as<double>(), 6 * static_cast<size_t>(std::max(npairs, 1)), 0.0,
                   ctx->stream));
  for (const QClass& hc : ctx->hclasses) {
    EriArgs a = base_args(ctx);
    a.grad_pair = dgp.as<double>();
    a.tiles = ctx->tiles_d.as<Tile>() + hc.tile_begin;
    a.ntiles = hc.ntiles;
    a.D = ctx->scf_D.as<double>();
    a.prim_alpha = ctx->alpha_d.as<double>();
    a.pair_atoms = dpa.as<int2>();
    a.grad = dgrad.as<double>();
    // one quartet per lane where the class's recurrences fit a lane's registers
    const EriGradFn lane_fn = ctx->force_generic ? nullptr
                                                 : eri_grad_launcher(hc.la, hc.lb, hc.lc, hc.ld);
    if (lane_fn) {
      HIPC(lane_fn(a, ctx->stream));
      continue;
    }
    const GenericProgram* gp = nullptr;
    TRY(find_program(ctx, hc.la, hc.lb, hc.lc, hc.ld, 1, &gp));
    if (gp) {
      HIPC(eri_generic_launch(a, *gp, kModeGrad, kBoysExact, 0, ctx->stream, ctx->lds_block, 0));
      continue;
    }
    // one program per centre
    for (int c = 0; c < 3; ++c) {
      TRY(find_program(ctx, hc.la, hc.lb, hc.lc, hc.ld, 2 + c, &gp));
      if (!gp)
        return ctx->fail(RCHEM_ERR_UNSUPPORTED, "gradient of ERI class " +
                                                    class_name(hc.la, hc.lb, hc.lc, hc.ld) +
                                                    " not served by this library");
      HIPC(eri_generic_launch(a, *gp, kModeGrad, kBoysExact, 0, ctx->stream, ctx->lds_block, 0));
    }
  }
  HIPC(grad_pairs_launch(dgp.as<double>(), dpa.as<int2>(), npairs, dgrad.as<double>(),
                         ctx->stream));
  // one-electron part
  std::vector<ShellDev> sh(ns);
  for (int s = 0; s < ns; ++s) {
    for (int x = 0; x < 3; ++x) sh[s].xyz[x] = ctx->shells[s].xyz[x];
    sh[s].l = ctx->shells[s].l;
    sh[s].nprim = ctx->shells[s].nprim;
    sh[s].p0 = ctx->shells[s].p0;
    sh[s].f0 = ctx->shells[s].f0;
  }
  std::vector<int2> pr;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) pr.push_back(make_int2(a, b));
  int gcls_begin[10];
  sort_pairs_by_class(sh, &pr, gcls_begin);
  HIPC(dsh.ensure(sizeof(ShellDev) * ns));
  HIPC(dex.ensure(sizeof(double) * ctx->exps.size()));
  HIPC(dcc.ensure(sizeof(double) * ctx->ccoef.size()));
  HIPC(dpr.ensure(sizeof(int2) * pr.size()));
  HIPC(dax.ensure(sizeof(double) * 3 * natom));
  HIPC(daz.ensure(sizeof(double) * natom));
  HIPC(dsat.ensure(sizeof(int) * ns));
  HIPC(hipMemcpyAsync(dsh.p, sh.data(), sizeof(ShellDev) * ns, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dex.p, ctx->exps.data(), sizeof(double) * ctx->exps.size(),
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dcc.p, ctx->ccoef.data(), sizeof(double) * ctx->ccoef.size(),
                      hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dpr.p, pr.data(), sizeof(int2) * pr.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(dax.p, xyz, sizeof(double) * 3 * natom, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(daz.p, Z, sizeof(double) * natom, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dsat.p, satom.data(), sizeof(int) * ns, hipMemcpyHostToDevice, ctx->stream));
  OneElGradArgs g{};
  g.shells = dsh.as<ShellDev>();
  g.exps = dex.as<double>();
  g.ccoef = dcc.as<double>();
  g.pairs = dpr.as<int2>();
  g.npairs = static_cast<int>(pr.size());
  g.nbf = n;
  g.natom = natom;
  g.atom_xyz = dax.as<double>();
  g.atom_z = daz.as<double>();
  g.boys_table = ctx->boys_d.as<double>();
  g.shell_atom = dsat.as<int>();
  g.D = ctx->scf_D.as<double>();
  g.W = dW.as<double>();
  g.grad = dgrad.as<double>();
  for (int c = 0; c < 10; ++c) g.cls_begin[c] = gcls_begin[c];

  HIPC(one_electron_grad_launch(g, ctx->stream));
  HIPC(hipMemcpyAsync(grad, dgrad.p, sizeof(double) * 3 * natom, hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  // nuclear repulsion: d/dR_A sum_{B<C} Z_B Z_C / R_BC
  for (int A = 0; A < natom; ++A)
    for (int B = 0; B < natom; ++B) {
      if (B == A) continue;
      double d[3], r2 = 0.0;
      for (int x = 0; x < 3; ++x) {
        d[x] = xyz[3 * A + x] - xyz[3 * B + x];
        r2 += d[x] * d[x];
      }
      const double r3 = r2 * std::sqrt(r2);
      for (int x = 0; x < 3; ++x) grad[3 * A + x] -= Z[A] * Z[B] * d[x] / r3;
    }
  return rc;
}

int rchem_scf(rchem_ctx* ctx, int natom, const double* Z, const double* xyz,
              const rchem_scf_opts* opts, rchem_scf_result* res, double* D_out) {
  const int rc = scf_run(ctx, natom, Z, xyz, opts, res, 0.0);
  if ((rc == RCHEM_OK || rc == RCHEM_ERR_NOT_CONVERGED) && D_out) {
    const size_t n2 = static_cast<size_t>(ctx->nbf) * ctx->nbf;
    HIPC(hipMemcpy(D_out, ctx->scf_D.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
  }
  return rc;
}

}  // extern "C"

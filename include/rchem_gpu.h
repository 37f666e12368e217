/*
 * rchem_gpu.h -- C ABI of the MI355X (gfx950) restricted-Hartree-Fock hot path.
 *
 * Drop-in boundary for berquist/rchem.  Every entry point below names the
 * reference interface it replaces (file:line in the rchem tree).  The ABI is
 * plain C: opaque context, plain pointers and sizes, int status codes.
 *
 * Conventions (fixed by rchem, reproduced exactly):
 *   - chemists' notation (mu nu|la si);
 *   - basis functions ordered atom -> basis-set shell -> l in the shell's
 *     angular_momentum list -> Cartesian component in get_ijk_list order
 *     (basis.rs:186-205, shell.rs:1-12), e.g. d = xx, xy, xz, yy, yz, zz;
 *   - primitive norm per Cartesian component (basis.rs:140-149), contraction
 *     coefficients used raw, no renormalisation (basis.rs:201);
 *   - matrices are n x n, row-major, double precision;
 *   - D = C_occ C_occ^T (no factor 2), F = H + 2J - K, E_elec = sum D o (H+F)
 *     (rchem.rs:101-127);
 *   - coordinates are used as given (bohr), no unit conversion (rchem.rs:33-35).
 *
 * Threading: a context owns one HIP stream on one device; calls on one
 * context must not overlap.  Distinct contexts are independent.
 * Memory: the library owns its device memory; callers own host buffers.
 * Errors: every function returns RCHEM_OK or an error code; the message of the
 * last failure on a context is rchem_last_error(ctx).  Nothing aborts.
 */
#ifndef RCHEM_GPU_H
#define RCHEM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rchem_ctx rchem_ctx;

enum rchem_status {
  RCHEM_OK = 0,
  RCHEM_ERR_INVALID = 1,       /* bad argument (null pointer, size, unknown name) */
  RCHEM_ERR_HIP = 2,           /* HIP runtime or library failure */
  RCHEM_ERR_NO_DEVICE = 3,     /* no usable gfx950 device */
  RCHEM_ERR_UNSUPPORTED = 4,   /* angular momentum or option not compiled in */
  RCHEM_ERR_NO_BASIS = 5,      /* operation needs a basis; none uploaded */
  RCHEM_ERR_NOT_CONVERGED = 6, /* SCF hit max_iter (results still written) */
  RCHEM_ERR_NONFINITE = 7,     /* NaN/Inf detected in an input or result */
  RCHEM_ERR_ALLOC = 8          /* device or host allocation failed */
};

/* Boys-function evaluation used for electron-repulsion integrals. */
enum rchem_boys_mode {
  RCHEM_BOYS_EXACT = 0,  /* accurate F_m(T) (production) */
  RCHEM_BOYS_COMPAT = 1  /* the reference's NR incomplete gamma, cints.c:302-373 */
};

enum rchem_precision {
  RCHEM_PREC_FP64 = 0,   /* every quartet in FP64 */
  RCHEM_PREC_MIXED = 1   /* quartets with Q_ab Q_cd max|D| < mixed_thresh in FP32 */
};

typedef struct rchem_opts {
  int boys_mode;         /* rchem_boys_mode, default RCHEM_BOYS_EXACT */
  int precision;         /* rchem_precision, default RCHEM_PREC_FP64 */
  double schwarz_thresh; /* skip quartets with Q_ab Q_cd below this; default 1e-12, 0 keeps all */
  double mixed_thresh;   /* FP32 cut for RCHEM_PREC_MIXED; default 1e-7 */
} rchem_opts;

/* Fills *o with the defaults above. */
int rchem_opts_default(rchem_opts *o);

/* ---- context -------------------------------------------------------------- */

/* Creates a context on HIP device `device` (index among visible devices).
 * No counterpart in rchem (single-threaded host code); the context holds what
 * rchem keeps in Basis (basis.rs:176-179) plus device state. */
int rchem_ctx_create(int device, rchem_ctx **out);
int rchem_ctx_destroy(rchem_ctx *ctx);
/* Message of the last failed call on ctx ("" if none).  Valid until the next
 * call on ctx.  With ctx == NULL, the last rchem_ctx_create failure. */
const char *rchem_last_error(const rchem_ctx *ctx);

/* ---- basis ---------------------------------------------------------------- */

/* Basis::new(atomnos, atomcoords, basis_set_name) (basis.rs:182-212), with the
 * basis-set data embedded (STO-3G, 6-31G for H, C, N, O) instead of fetched
 * from basis_set_exchange (basis.rs:87-106).  xyz is natom x 3 (bohr). */
int rchem_basis_from_atoms(rchem_ctx *ctx, int natom, const int *Z, const double *xyz,
                           const char *basis_name);

/* The same from explicit shells: shell s has centre xyz[3s..3s+2], angular
 * momentum l[s] (0..2), nprim[s] (1..8) primitives whose exponents and raw contraction
 * coefficients are exps/coefs[off_s .. off_s+nprim[s]) with off_s the running
 * sum of nprim.  Functions are numbered shell by shell, components in
 * get_ijk_list order. */
int rchem_basis_upload(rchem_ctx *ctx, int nshell, const double *xyz, const int *l,
                       const int *nprim, const double *exps, const double *coefs);

/* Number of Cartesian basis functions (rchem: basis_set.cgtos.len()). */
int rchem_basis_nbf(const rchem_ctx *ctx, int *nbf);
/* Number of shells and total primitives of the current basis. */
int rchem_basis_dims(const rchem_ctx *ctx, int *nshell, int *nprim_total);
/* Copies the current shells out (arrays sized by rchem_basis_dims). */
int rchem_basis_shells(const rchem_ctx *ctx, double *xyz, int *l, int *nprim, double *exps,
                       double *coefs);

/* ---- one-electron integrals --------------------------------------------- */

/* S, T and V = sum_C Z_C <mu|-1/r_C|nu> (basis.rs:234, :273, :316 and
 * os86.rs:629-773), n x n each, host pointers; any of S/T/V may be NULL. */
int rchem_one_electron(rchem_ctx *ctx, int natom, const double *Z, const double *xyz,
                       double *S, double *T, double *V);

/* ---- two-electron integrals --------------------------------------------- */

/* JK_direct(J, K, basis, D) (basis.rs:383-428): J_mn = sum (mn|ls) D_ls,
 * K_mn = sum (ml|ns) D_ls; J and K are overwritten.  D must be symmetric.
 * Host pointers, n x n.  opts may be NULL (defaults).  The reference loops over
 * every primitive quartet; this uses Schwarz screening and 8-fold symmetry. */
int rchem_jk(rchem_ctx *ctx, const double *D, double *J, double *K, const rchem_opts *opts);

/* Same with device pointers (HBM-resident D, J, K; n x n doubles each).
 * Runs on the context's stream and returns without synchronising, so D is
 * not inspected on the host: a non-finite entry of D propagates into J and K
 * (never screened away), where rchem_jk returns RCHEM_ERR_NONFINITE. */
int rchem_jk_device(rchem_ctx *ctx, const double *dD, double *dJ, double *dK,
                    const rchem_opts *opts);

/* The same build without allocating: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in rchem.  The build's accumulators (the
 * lower-half J and K, the pair-indexed Coulomb accumulators and density
 * blocks, the planner's statistics) live in a fixed 48 MiB device array that
 * the library holds on every GPU it runs on until the process ends; the call
 * allocates and frees nothing.  It uses the context's plan for
 * opts->schwarz_thresh, which the first rchem_jk / rchem_jk_device call with
 * the same threshold builds (that call allocates): without one, or when the
 * basis needs more than 48 MiB of accumulators (about 3 n^2 doubles: n up to
 * roughly 1400 functions), it returns RCHEM_ERR_INVALID and launches nothing.
 * Every call on one GPU shares the array, whatever its context, and keeps the
 * last call's contents: calls that overlap on one GPU (two streams, two
 * threads, two contexts) race and must be serialised by the caller.  Per-class
 * timing is not recorded.  Same results as rchem_jk_device. */
int rchem_jk_device_noalloc(rchem_ctx *ctx, const double *dD, double *dJ, double *dK,
                            const rchem_opts *opts);

/* build_I(basis) (basis.rs:430-460): the full n^4 tensor [mu,nu,la,si]
 * row-major into host memory (n^4 doubles).  No screening. */
int rchem_eri_tensor(rchem_ctx *ctx, double *I, const rchem_opts *opts);

/* Selected shell quartets: for quartet q, shells[4q..4q+3] = (A, B, C, D) of
 * the current basis; out receives nA nB nC nD values per quartet, quartets
 * back to back, each block row-major over (a, b, c, d) components in rchem's
 * order.  No screening.  No rchem counterpart (rchem has no shell-level API):
 * it samples integrals of systems too large for the full tensor. */
int rchem_eri_shell_quartets(rchem_ctx *ctx, int nq, const int *shells, double *out,
                             const rchem_opts *opts);

/* Schwarz bounds sqrt(max_comp |(ab|ab)|) for every shell pair a >= b, as an
 * nshell x nshell lower-triangular-filled matrix (host).  No rchem
 * counterpart (rchem does not screen). */
int rchem_schwarz(rchem_ctx *ctx, double *Q, const rchem_opts *opts);

/* Statistics of the last rchem_jk / rchem_jk_device call. */
typedef struct rchem_jk_stats {
  double quartets;           /* unique contracted shell quartets computed */
  double primitive_quartets; /* primitive quartets computed */
  double integrals;          /* contracted Cartesian integrals computed */
  double flops;              /* algorithmic FP operations (FMA = 2) */
  double quartets_fp32;      /* of `quartets`, computed in FP32 */
  double tiles;              /* wavefront tiles (64 lanes) the quartets were packed into */
} rchem_jk_stats;
int rchem_jk_last_stats(rchem_ctx *ctx, rchem_jk_stats *st);

/* Per-class kernel timing.  With profiling on, every rchem_jk / rchem_jk_device
 * call records a HIP event pair around each (la lb|lc ld) class's launch on the
 * context stream; rchem_jk_class_timings sums them since the last reset.
 * No rchem counterpart (rchem has no timing). */
typedef struct rchem_class_timing {
  int la, lb, lc, ld;
  int fp32;                  /* 1: the FP32 kernel of a mixed-precision build */
  int launches;              /* launches summed into ms */
  double ms;                 /* summed kernel time */
  double quartets;           /* per launch */
  double primitive_quartets; /* per launch */
  double flops;              /* per launch, algorithmic (FMA = 2) */
} rchem_class_timing;
int rchem_set_profiling(rchem_ctx *ctx, int enable);
int rchem_reset_class_timings(rchem_ctx *ctx);
/* Writes up to max entries; *count receives the number of classes timed. */
int rchem_jk_class_timings(rchem_ctx *ctx, rchem_class_timing *out, int max, int *count);

/* Launch tuning of the Fock build (no rchem counterpart; rchem has no GPU).
 * Each (la lb|lc ld) class and precision (fp32 = 1: the FP32 kernels of a
 * mixed build) has a launch shape of its LDS-digestion kernel (an index into
 * the library's compiled shapes, rchem_lds_variant_info), a persistent
 * workgroup count and a chunk size in 64-lane tiles.  A context starts from the
 * table data/tuning_gfx950.txt beside the library (RCHEM_TUNING_FILE names
 * another; empty: none), written by rchem_amd/autotune.py; classes without an
 * entry use shape 0, 2048 workgroups, 64 tiles.  Lines of the table:
 * "la lb lc ld fp32 variant blocks chunk_tiles [ket_order]", '#' comments.  A shape a class
 * is not compiled in falls back to shape 0.  Setting an entry re-plans the
 * next build (which allocates). */
int rchem_lds_variant_info(int variant, int *waves, int *min_waves_per_simd, int *prefetch,
                           int *unroll, int *branch_free);
int rchem_set_class_tuning(rchem_ctx *ctx, int la, int lb, int lc, int ld, int fp32, int variant,
                           int blocks, int chunk_tiles);
int rchem_get_class_tuning(const rchem_ctx *ctx, int la, int lb, int lc, int ld, int fp32,
                           int *variant, int *blocks, int *chunk_tiles);
/* LDS chunks of a class and precision in ket order (ket_order 1; 0: in bra
 * order; -1: the default, FP64 lists in ket order, FP32 lists in bra order):
 * the planner sorts the chunks by their first ket pair, so that the
 * workgroups running at one time read the same ket records.  The table's
 * optional ninth column sets it; RCHEM_CHUNK_ORDER=0 / 1 overrides every
 * class.  get reports the order in effect. */
int rchem_set_class_ket_order(rchem_ctx *ctx, int la, int lb, int lc, int ld, int fp32,
                              int ket_order);
int rchem_get_class_ket_order(const rchem_ctx *ctx, int la, int lb, int lc, int ld, int fp32,
                              int *ket_order);
/* The batched SCF's table: rchem_batch_scf uses it in place of the context's
 * while it runs (multi-molecule bases of small molecules favour other launch
 * shapes).  A context loads data/tuning_gfx950_batch.txt beside the library
 * (RCHEM_TUNING_BATCH_FILE names another; empty: none), written by
 * `python -m rchem_amd.autotune --workload batch512`; without one the batch
 * uses the context's table.  NULL or "": none; *entries (may be NULL)
 * receives the count. */
int rchem_load_batch_tuning(rchem_ctx *ctx, const char *path, int *entries);
/* Replaces the context's table by the file at path (NULL or "": no table,
 * every class at its defaults); *entries (may be NULL) receives the count. */
int rchem_load_tuning(rchem_ctx *ctx, const char *path, int *entries);

/* ---- SCF ------------------------------------------------------------------ */

typedef struct rchem_scf_opts {
  int nocc;            /* doubly occupied orbitals (rchem.rs:59,78) */
  int max_iter;        /* default 1024 (rchem.rs:65) */
  double thresh_e;     /* |dE_elec| convergence, default 1e-11 (rchem.rs:63) */
  rchem_opts jk;       /* integral options */
  int verbose;         /* 1: print rchem's iteration line (rchem.rs:83) */
  int diis;            /* 0 (default): rchem's plain iterations; k > 0: Pulay DIIS over
                          the last k Fock matrices with the orthogonalised commutator
                          X^T (FDS - SDF) X as error, converged when also
                          max |error| < 1e-8 (rchem has no DIIS, rchem.rs:72-90) */
  int graph;           /* 1: replay iterations from hipGraphs.  With diis > 0 and
                          purify = 1 the whole iteration -- Fock build, energy,
                          commutator, DIIS products and solve, extrapolation,
                          purification, density, damping -- runs on the device with
                          no host step and is captured once (after one eager
                          iteration) and replayed, one host read of a few status
                          scalars per iteration; otherwise the Fock build (planner,
                          ERI classes, digestion, F = H + 2J - K) alone.  Recaptured
                          at the precision switch; a capture the runtime refuses
                          falls back to kernel-by-kernel launches.  0 (default):
                          kernel by kernel */
  int purify;          /* 1: each iteration's density D = X P X^T with P, the projector
                          onto the nocc lowest eigenvectors of F' = X^T F X, from
                          canonical purification (GEMMs only) instead of the
                          eigensolver, which takes over only an iteration whose F'
                          has no gap.  Single-molecule SCF and gradient (whose
                          energy-weighted density is W = D F D either way); the
                          batched SCF ignores it.  0 (default): rchem's eigh every
                          iteration (rchem.rs:74-78) */
  double damping;      /* a in [0, 1): each new density is mixed with the previous one,
                          D <- (1 - a) D_new + a D_old (a convergence aid for hard
                          cases; the converged density is the same stationary point);
                          0 (default): rchem's undamped update */
  double damping_until; /* damping applies while the orthogonalised commutator
                          max |X^T (FDS - SDF) X| (with DIIS) or |dE| (without) is at
                          least this, per molecule in a batch; 0: every iteration */
} rchem_scf_opts;

typedef struct rchem_scf_result {
  double e_elec;   /* sum D o (H + F) (rchem.rs:106-108, :80) */
  double e_nuc;    /* sum Z_A Z_B / R_AB (not printed by rchem) */
  double e_total;
  int iterations;
  int converged;
  int fp64_from;   /* mixed precision: first iteration with all-FP64 J/K (the switch
                      comes when |dE| < 1e-7 Eh; in a batched SCF per molecule, at
                      iteration 64 at the latest); 0 for FP64 runs, -1 if never */
  int graph_iterations; /* iterations replayed from a whole-iteration hipGraph
                      (opts->graph with DIIS and purification); 0 otherwise */
} rchem_scf_result;

int rchem_scf_opts_default(rchem_scf_opts *o);

/* rchem's main SCF (rchem.rs:40-90) on the current basis for the given
 * nuclei: Loewdin X, core-Hamiltonian guess, plain iterations until
 * |dE_elec| < thresh_e.  D_out (n x n, may be NULL) receives the final
 * density.  Returns RCHEM_ERR_NOT_CONVERGED if max_iter is reached. */
int rchem_scf(rchem_ctx *ctx, int natom, const double *Z, const double *xyz,
              const rchem_scf_opts *opts, rchem_scf_result *res, double *D_out);

/* Batched SCF (no counterpart in rchem, which solves one molecule): nmol
 * independent closed-shell molecules, molecule m with natom[m] atoms whose
 * atomic numbers and coordinates (bohr) follow those of molecule m-1 in Z and
 * xyz, all in basis basis_name.  Molecules are grouped into chunks of at most
 * 8192 basis functions; a chunk is one multi-molecule basis whose Fock builds
 * form only quartets within a molecule (one launch per class for every
 * molecule of the chunk), and whose SCF step runs batched over the molecules
 * on the device (strided-batched rocSOLVER eigensolves, per-molecule DIIS when
 * opts->diis > 0, else rchem's plain iterations).  opts->nocc is ignored (each
 * molecule has half its electron count); in mixed precision each molecule
 * switches to all-FP64 quartets once its energy changes by less than 1e-7 Eh
 * per iteration (at iteration 64 at the latest).  A converged molecule retires
 * from the Fock builds; the chunk iterates until every molecule has converged
 * or max_iter.  res receives nmol results; converged, iterations and fp64_from
 * are per molecule.  opts->graph replays each chunk's Fock build as a hipGraph;
 * opts->purify is ignored.  The context's basis is cleared afterwards. */
int rchem_batch_scf(rchem_ctx *ctx, int nmol, const int *natom, const int *Z,
                    const double *xyz, const char *basis_name,
                    const rchem_scf_opts *opts, rchem_scf_result *res);

/* Analytic RHF nuclear gradient on the current basis (rchem has no derivative
 * code; the energy it differentiates is rchem's, rchem.rs:72-90, plus the
 * nuclear repulsion).  Runs rchem_scf's iterations until |dE_elec| <
 * opts->thresh_e and the density changes by less than 1e-9 per iteration,
 * then
 *   dE/dR = 2 sum D dH + sum Gamma d(mu nu|la si) - 2 sum W dS + dE_nuc/dR,
 *   Gamma = 2 D_mn D_ls - (D_ml D_ns + D_ms D_nl) / 2,  W = C_occ eps_occ C_occ^T,
 * with the derivative integrals recomputed on the device (no quartet is
 * stored).  grad: 3 x natom doubles, Eh/bohr, atom-major.  Every shell centre
 * must coincide with one of the nuclei.  res as for rchem_scf.  Every class
 * with l <= 2 is served (s, p and d shells).  If the SCF does not converge in
 * opts->max_iter iterations the call returns RCHEM_ERR_NOT_CONVERGED, fills
 * res, and leaves grad untouched (the formula holds at a stationary density
 * only). */
int rchem_rhf_gradient(rchem_ctx *ctx, int natom, const double *Z, const double *xyz,
                       const rchem_scf_opts *opts, rchem_scf_result *res, double *grad);

#ifdef __cplusplus
}
#endif

#endif /* RCHEM_GPU_H */

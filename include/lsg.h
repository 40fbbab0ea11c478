/*
 * lsg.h - C ABI of the MI355X-native DAFMM+NCA Lippmann-Schwinger solver.
 *
 * This is the drop-in boundary for the reference's matvec/GMRES path
 * (vaishna77/Lippmann_Schwinger_Solver, GMRES/ project):
 *
 *   reference                                              replaced by
 *   ---------------------------------------------------    ------------------------
 *   struct inputsToDFMM + argv parsing  GMRES/GMRES.cpp:129-138,322-346   lsg_inputs
 *   new DFMM(inputs)                    GMRES/GMRES.cpp:152-157           lsg_create
 *   DFMM::FMMTreeInitialisation()       GMRES/GMRES.cpp:161-199           lsg_create
 *   DFMM::MatVecProduct(Vec&, Vec&)     GMRES/GMRES.cpp:201-254           lsg_matvec
 *   RestartedGMRES<DFMM>(...)           GMRES/GMRES.cpp:307-319           lsg_gmres
 *   classGMRES::gmres (tolerance stop)  GMRES_Hybrid/include/gmres.hpp:53-119  lsg_gmres
 *   userkernel::RHSFunction             GMRES/GMRES.cpp:122-125           lsg_rhs
 *   userkernel::IncidenceFunction       GMRES/GMRES.cpp:118-121           lsg_incident
 *   DFMM::gridPoints                    GMRES/GMRES.cpp:145,171           lsg_grid_points
 *   delete S                            GMRES/GMRES.cpp:648               lsg_destroy
 *
 * Conventions
 *   - Complex vectors are interleaved complex128 (re, im) double pairs, the
 *     layout of Eigen::VectorXcd; length N = (#leaves) * nChebNodes^2, unknown
 *     index = leafIndex * nChebNodes^2 + (iy * nChebNodes + ix).
 *   - The caller owns every host buffer; the context owns device memory.
 *   - Every call is synchronous. One context must not be used from two host
 *     threads at once (the reference tree is not re-entrant either).
 *   - Return value 0 = success, negative = error; lsg_last_error() gives text.
 *     There is no silent CPU fallback: without a usable GPU lsg_create fails.
 */
#ifndef LSG_H_
#define LSG_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSG_OK 0
#define LSG_ERR_ARG -1       /* invalid argument */
#define LSG_ERR_DEVICE -2    /* HIP runtime / device error */
#define LSG_ERR_SETUP -3     /* tree / table / NCA construction failed */
#define LSG_ERR_NUMERIC -4   /* NaN/Inf input, zero right-hand side, breakdown */
#define LSG_ERR_NOCONV -5    /* GMRES hit its iteration cap above tolerance */

#define LSG_MAX_BUMPS 16

/* Contrast q(x). kind 0..5 are the reference's Qchoice functions
 * (gaussian, multipleGaussians, plasma, flatBump, cavity, lens;
 * GMRES/GMRES.cpp:5-117). kind 6 is a sum of Gaussian bumps
 * q(x) = sum_i amp_i * exp(-|x - c_i|^2 / width_i^2). */
typedef struct lsg_contrast {
  int kind;
  int n_bumps;
  double cx[LSG_MAX_BUMPS];
  double cy[LSG_MAX_BUMPS];
  double width[LSG_MAX_BUMPS];
  double amp[LSG_MAX_BUMPS];
} lsg_contrast;

/* The reference's 12 run-time inputs (GMRES/GMRES.cpp:322-333) plus the
 * choices the reference makes by editing source. */
typedef struct lsg_inputs {
  int nCones_LFR;      /* argv[1]  cones per box at level level_LFR-1 */
  int nChebNodes;      /* argv[2]  Chebyshev nodes per leaf side */
  int treeAdaptivity;  /* argv[3]  adaptive-tree tolerance exponent */
  double L;            /* argv[4]  half side of the domain [-L,L]^2 */
  double kappa;        /* argv[5]  wavenumber */
  int yes2DFMM;        /* argv[6]  1: directional (HFR) levels enabled */
  int degreeOfBases;   /* argv[7]  total-degree bound of the interpolation basis */
  int TOL_POW;         /* argv[8]  NCA tolerance 10^-TOL_POW */
  int m;               /* argv[9]  Krylov dimension (names the result dir) */
  int restart;         /* argv[10] GMRES restarts */
  int nLevelsUniform;  /* argv[11] depth of the uniform tree */
  int Qchoice;         /* argv[12] contrast selector (== contrast.kind) */
  int tree_mode;       /* 0: uniform tree (createUniformTree); 1: adaptive */
  int find_phi;        /* 1: A = I + diag(k^2 q) V (solve); 0: V (field) */
  double inc_angle;    /* plane-wave direction; 0 = +x as in the reference */
  lsg_contrast contrast;
  int setup_device;    /* 0: NCA skeletons and operator entries on the host (skeletons
                          identical to the reference algorithm's); 1: on the GPU (batched ACA
                          and entry blocks; pivots agree with the host's except where two
                          candidates tie to rounding) */
  const char* precomputations_dir; /* NULL: compute the far-field and near-field tables;
                          else read them from <dir>/precomputations/M and /Neighbor in the
                          reference's cache layout and text format (G_FMM2DTree.hpp:1301-1313,
                          1911-1950); a missing or malformed file fails lsg_create */
  int aca_exit_pow;    /* 0: the reference's absolute exit of the ACA on the new cross vectors'
                          norms, 1e-8 (GMRES/ACA.hpp:350, 453); p in [8, 20]: 10^-p instead. That
                          fixed exit, not TOL_POW, bounds how close the DAFMM comes to the dense
                          operator (5.7e-6 at depth 3 for any TOL_POW); with TOL_POW 14 and p 16 the
                          DAFMM matches the dense matvec to ~1e-12 */
} lsg_inputs;

/* Work and size figures of a built operator (for roofline accounting). */
typedef struct lsg_info {
  int64_t N;              /* unknowns */
  int nLevels;            /* leaf level of the tree */
  int level_LFR;          /* first low-frequency level */
  int n_leaves;
  int64_t E_far;          /* complex entries of the fused far-field operators */
  int64_t E_near;         /* near-field complex MACs per matvec */
  int64_t bytes_matvec;   /* algorithmic HBM bytes of one FP64 matvec */
  int64_t flops_matvec;   /* algorithmic flops of one matvec */
  int64_t n_state;        /* complex slots in the device state vector */
  int max_rank;           /* largest NCA rank */
  double setup_seconds;   /* host operator build time */
  double setup_phase_seconds[4]; /* of which: far-field tables, NCA skeletons, operator fill, near-field tables */
} lsg_info;

/* GMRES statistics. */
typedef struct lsg_stats {
  int iterations;         /* Arnoldi steps (= matvecs excluding residual checks) */
  int restarts;           /* completed restart cycles */
  int converged;          /* 1 if the true FP64 residual met tol */
  double rel_residual;    /* true ||b - A x|| / ||b|| in FP64 */
  double est_residual;    /* Givens residual estimate at exit */
  double seconds;         /* wall time */
  int refinements;        /* FP64 refinement steps (LSG_SOLVE_FP32_IR) or outer steps (FP32_FGMRES), else 0 */
} lsg_stats;

/* Solver selection for lsg_gmres_ex. */
#define LSG_SOLVE_FP64 0     /* GMRES with the FP64 operator (lsg_gmres) */
#define LSG_SOLVE_FP32_IR 1  /* FP64 iterative refinement around GMRES with the FP32 operator */
#define LSG_SOLVE_FP32_FGMRES 2 /* lsg_gmres_batch only: FP64 flexible GMRES(max_outer) whose preconditioner is
                                   one FP32 GMRES(restart_len) solve to inner_tol per outer step */

typedef struct lsg_solve_opts {
  int mode;               /* LSG_SOLVE_FP64, LSG_SOLVE_FP32_IR or (batched) LSG_SOLVE_FP32_FGMRES */
  int restart_len;        /* Krylov dimension per cycle, 1..500 */
  int max_restarts;       /* cycle cap (per inner solve in FP32_IR mode) */
  double tol;             /* target true FP64 relative residual (> 0 for FP32_IR) */
  double inner_tol;       /* FP32_IR: floor of the inner relative residual (e.g. 1e-6); FP32_FGMRES: inner tolerance */
  int max_outer;          /* FP32_IR: refinement cap; FP32_FGMRES: outer Krylov dimension per cycle (<= 500) */
  int device_ptrs;        /* 1: b and x are device addresses of N complex128 */
} lsg_solve_opts;

typedef struct lsg_ctx lsg_ctx;

/* Fill a kind-6 contrast with n seeded bumps (SplitMix64 stream, draws per
 * bump in the order cx, cy, width, amp, each uniform in its range). */
int lsg_bumps_from_seed(uint64_t seed, int n, double cmin, double cmax,
                        double wmin, double wmax, double amin, double amax,
                        lsg_contrast* out);

/* Build tree, tables, NCA skeletons and fused operators on the host, upload
 * them to the current HIP device. Replaces new DFMM + FMMTreeInitialisation. */
int lsg_create(const lsg_inputs* in, lsg_ctx** out);
void lsg_destroy(lsg_ctx* ctx);
const char* lsg_last_error(const lsg_ctx* ctx);
/* Error text of the last failed lsg_create (no context exists then). */
const char* lsg_create_error(void);

int lsg_get_info(const lsg_ctx* ctx, lsg_info* info);
/* Leaf-major node coordinates, x[N], y[N]. */
int lsg_grid_points(const lsg_ctx* ctx, double* x, double* y);
/* Leaf boxes in leaf order: centres and half side (collectBoundary input,
 * GMRES/G_FMM2DTree.hpp:1165-1177). Arrays of n_leaves. */
int lsg_leaf_boxes(const lsg_ctx* ctx, double* cx, double* cy, double* radius);
/* k^2 q(x_i) at every node (the diagonal D of A = I + D V). */
int lsg_contrast_at_nodes(const lsg_ctx* ctx, double* qk2);
/* u_inc(x_i) = exp(i k (x cos a + y sin a)), interleaved complex. */
int lsg_incident(const lsg_ctx* ctx, double* u);
/* b_i = -k^2 q(x_i) u_inc(x_i) (RHSFunction). */
int lsg_rhs(const lsg_ctx* ctx, double* b);

/* Re-interpolation of a node field f (N complex, leaf-major as every vector of
 * this API) onto npts points (x[i], y[i]): per leaf the least-squares
 * coefficients c = Q_pinv f_leaf of the total-degree Chebyshev basis
 * T_a(x) T_b(y), a + b < degreeOfBases, evaluated at each point in the leaf box
 * [centre - r, centre + r) holding it. Restates the Phi -> Psi step of
 * GMRES_Hybrid/LippmannSchwinger/main/hybridTrueValidation_gmres.cpp:186-209
 * (the field of one tree at the nodes of another). Host arrays; out: npts
 * complex. A point outside [-L, L)^2 gives LSG_ERR_ARG (the reference falls
 * back to the last leaf there). Uniform trees. */
int lsg_interpolate(lsg_ctx* ctx, const double* f, const double* x, const double* y, int64_t npts, double* out);

/* Writes the table cache of these inputs under <dir>/precomputations in the
 * reference's layout and text format: far-field tables M of levels
 * 0..nLevelsUniform (writeMToFile, G_FMM2DTree.hpp:1252-1269) and the 53
 * neighbour tables of levels 2..10 (writeNeighborInteractionsForLevel,
 * :1813-1909). Host only; errors in lsg_create_error(). */
int lsg_write_precomputations(const lsg_inputs* in, const char* dir);

/* y = A x for nrhs right-hand sides (column j at x + 2*ld*j doubles),
 * FP64 on the device. Replaces DFMM::MatVecProduct. nrhs >= 2 runs the block
 * operator (64 right-hand sides per FP64 MFMA block matvec). */
int lsg_matvec(lsg_ctx* ctx, const double* x, double* y, int nrhs, int64_t ld);

/* Device-pointer variant: x, y are device addresses of N complex128 each;
 * the call is enqueued on the context's stream and returns without waiting. */
int lsg_matvec_device(lsg_ctx* ctx, const void* x_dev, void* y_dev);
/* Wait for all work queued on the context's stream. */
int lsg_sync(lsg_ctx* ctx);
/* Average device time (ms) of each matvec pass recorded with HIP events over
 * `reps` matvecs of a device-resident vector; writes up to max_pass values
 * into ms[] and the pass algorithmic bytes into bytes[]; returns #passes. */
int lsg_profile_passes(lsg_ctx* ctx, int reps, double* ms, int64_t* bytes,
                       int max_pass);

/* Average device time (ms) of one full matvec (all passes, both streams) over
 * `reps` matvecs, measured with HIP events on the context's stream. */
int lsg_time_matvec(lsg_ctx* ctx, int reps, double* ms);

/* Restarted GMRES on the device: x <- solution of A x = b (x holds the
 * initial guess). restart_len = Krylov dimension per cycle, max_restarts =
 * cycle cap. tol > 0: stop when the true FP64 relative residual <= tol;
 * tol <= 0: reference mode, run exactly max_restarts cycles of restart_len
 * Arnoldi steps (GMRES/GMRES.cpp:257-319, with its 1e-8 breakdown exit). */
int lsg_gmres(lsg_ctx* ctx, const double* b, double* x, int restart_len,
              int max_restarts, double tol, lsg_stats* stats);

/* Mixed-precision variant of lsg_matvec (no reference counterpart): the
 * operators and the accumulation inside every pass are FP32 (interleaved
 * re/im), x and y stay FP64; the FP32 copy of the operators is made on the
 * device on first use. nrhs >= 2 runs the low-precision block operator of the
 * inner solves (64 right-hand sides per pass, FP32 pass state; split-BF16 or
 * FP32 MFMA as lsg_set_inner_arith selects). */
int lsg_matvec_fp32(lsg_ctx* ctx, const double* x, double* y, int nrhs, int64_t ld);
/* lsg_time_matvec for the FP32 operator. */
int lsg_time_matvec_fp32(lsg_ctx* ctx, int reps, double* ms);
/* Contrast override for a volume-potential context (lsg_inputs.find_phi = 0,
 * whose compression does not depend on q): afterwards every matvec returns
 * x + qk2 .* (V x), the Lippmann-Schwinger operator I + diag(k^2 q) V of the
 * reference (G_FMM2DTree.hpp:2884-2960 with findPhi) for the contrast given
 * here. qk2 holds ncol columns of N values k^2 q(x_i) (column-major): the
 * single-vector calls use column 0; block calls and lsg_gmres_batch use column
 * q for right-hand side q (column 0 when q >= ncol). qk2 = NULL removes it.
 * No reference counterpart: the reference rebuilds the tree per contrast. */
int lsg_set_contrast(lsg_ctx* ctx, const double* qk2, int ncol);

/* Average device time (ms) of one block matvec (64 right-hand sides, FP64
 * MFMA) over `reps` products of a device-resident block. */
int lsg_time_matvec_block(lsg_ctx* ctx, int reps, double* ms);
/* Same for the FP32 block operator (FP32 operators and FP32 MFMA). */
int lsg_time_matvec_block_fp32(lsg_ctx* ctx, int reps, double* ms);

/* GMRES with solver options: lsg_gmres (mode LSG_SOLVE_FP64), or FP64
 * iterative refinement (mode LSG_SOLVE_FP32_IR): r = b - A x with the FP64
 * operator, d ~ A32^-1 r by GMRES with the FP32 operator (Gram-Schmidt and
 * Givens in FP64) to relative residual max(inner_tol, tol*||b||/(2||r||)),
 * x += d, until ||b - A x|| <= tol ||b|| in FP64. The refinement follows the
 * reference's GMRES cycle semantics (GMRES/GMRES.cpp:257-319) inside each
 * inner solve; the refinement loop itself has no reference counterpart. */
int lsg_gmres_ex(lsg_ctx* ctx, const double* b, double* x, const lsg_solve_opts* opts, lsg_stats* stats);

/* Allocation-free variant of lsg_gmres_ex (mode LSG_SOLVE_FP64, device
 * pointers). No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops).
 * b_dev and x_dev are device addresses of N complex128 (x holds the initial
 * guess); the Krylov workspace comes from one scratch array of 128 MiB per
 * device, held by the library until the process ends, and nothing is
 * allocated or freed. restart_len * N is bounded by that size (about 8.3 M
 * complex128 in total: GMRES(50) up to N = 161,000); larger requests fail
 * with LSG_ERR_ARG. Every call on one device shares its array, so calls that
 * overlap on one device (from two host threads or two contexts) race. Same
 * iterates and statistics as lsg_gmres_ex. Synchronous. */
int lsg_gmres_noalloc(lsg_ctx* ctx, const void* b_dev, void* x_dev, int restart_len, int max_restarts, double tol,
                      lsg_stats* stats);

/* Batched GMRES for nrhs <= 64 systems A x_q = b_q sharing the operator (the
 * reference solves one right-hand side per RestartedGMRES call,
 * GMRES/GMRES.cpp:307-319): b and x hold nrhs columns at b + 2*ld*q (host, or
 * device addresses when opts->device_ptrs), x the initial guesses. The systems
 * advance in lockstep through one block matvec (FP64 MFMA) per Arnoldi step,
 * each with its own Krylov basis (classical Gram-Schmidt), Givens rotations and
 * stopping point; each cycle opens with the true FP64 residual of every
 * system. opts->mode LSG_SOLVE_FP32_IR refines the batch in FP64 around
 * batched GMRES with the FP32 block operator (FP32 MFMA), each system to
 * max(inner_tol, tol ||b|| / (2 ||r||)). LSG_SOLVE_FP32_FGMRES runs FP64
 * flexible GMRES(max_outer) on the batch (at most 20 cycles), each outer step
 * preconditioned by one FP32 batched GMRES(restart_len) solve to inner_tol:
 * the corrections are re-weighted by the outer least-squares problem instead
 * of added. rel_residuals (host, nrhs, may be
 * NULL) receives the final true relative residuals; stats->rel_residual is
 * their maximum and stats->converged the number of systems at tol. */
int lsg_gmres_batch(lsg_ctx* ctx, const double* b, double* x, int nrhs, int64_t ld, const lsg_solve_opts* opts,
                    double* rel_residuals, lsg_stats* stats);

/* Autotuner (no reference counterpart): for every pass, time the one-wavefront,
 * one-workgroup-per-destination and one-lane-per-row kernels, the first upward
 * passes fused per subtree against their separate launches, and for the near field of a
 * uniform tree the FP64 MFMA kernel and the FMA kernel with 1 or 2 leaves per
 * thread, on this operator with HIP events (best of `reps`); keep the
 * fastest. n_changed (may be NULL) receives the number of changed choices;
 * lsg_tuning() returns one line per decision, followed by the block
 * operators' timed choices once they are enabled (tiles per wavefront, the
 * near field beside or after the passes); the string stays valid until the
 * next lsg_tuning() call on ctx. */
int lsg_autotune(lsg_ctx* ctx, int reps, int* n_changed);
const char* lsg_tuning(const lsg_ctx* ctx);

/* Pass kernel choice without timing (no reference counterpart): every pass of
 * the single-vector operator runs the given kernel, 0 one workgroup per
 * destination, 1 one wavefront per destination, 2 one lane per output row
 * (passes whose rows exceed 256 keep the workgroup kernel); 3, 4 and 5 are
 * the same with the leaf level's upward pass and the next two, and the last
 * two small downward passes, each group fused into one launch of one
 * workgroup per group of subtrees (where the operator has such groups). For
 * tests and measurements of each kernel; lsg_autotune replaces the choice. */
int lsg_set_pass_kernel(lsg_ctx* ctx, int kernel);

/* Right-hand-side tiles of 16 per wavefront in the block passes of the inner
 * solves' arithmetic (lsg_set_inner_arith; no reference counterpart): 1 (four
 * wavefronts per workgroup) or 2 (two, each operator value feeding two
 * MFMAs); in split BF16 on a uniform tree also 3 or 4 (1 or 2 tiles, the near
 * field computed on the second stream beside the upward and downward passes,
 * the leaf passes added after them). Enabling that block operator picks the
 * fastest by timing; this replaces that choice (tests, measurements). */
int lsg_set_block_tiles(lsg_ctx* ctx, int tiles);

/* Arithmetic of the FP32-inner batched solves (LSG_SOLVE_FP32_IR and
 * LSG_SOLVE_FP32_FGMRES in lsg_gmres_batch), of lsg_matvec_fp32 with nrhs >= 2
 * and of lsg_time_matvec_block_fp32 (no reference counterpart). Both keep the
 * pass state in FP32 and the inner Krylov basis in FP16:
 *   LSG_ARITH_SPLIT_BF16 (default): operators and source values split as
 *     a = ah + al (bf16 pair, 16 significant bits), products (ah + al)(bh + bl)
 *     as four BF16 MFMA terms with FP32 sums;
 *   LSG_ARITH_FP32: FP32 operators on FP32 MFMA, 24 significant bits. */
#define LSG_ARITH_FP32 0
#define LSG_ARITH_SPLIT_BF16 1
int lsg_set_inner_arith(lsg_ctx* ctx, int arith);

/* on = 1: capture the FP64 matvec (both streams) into a hipGraph, replayed by
 * every later FP64 matvec of the context (one launch plus a pointer update);
 * on = 0: back to individual launches. */
int lsg_use_graph(lsg_ctx* ctx, int on);

/* Ragged batch: n independent systems, each with its own context (tree depth,
 * wavenumber and contrast may all differ), solved concurrently by
 * lsg_gmres_ex with the same options, one host thread and one pair of streams
 * per context (matvecs replayed from each context's graph when lsg_use_graph
 * is on). b[i], x[i] are the vectors of context i; stats (may be NULL) has n
 * entries. Returns the first error other than LSG_ERR_NOCONV, else
 * LSG_ERR_NOCONV if any system missed tol, else LSG_OK. */
int lsg_gmres_ragged(lsg_ctx* const* ctxs, int n, const double* const* b, double* const* x,
                     const lsg_solve_opts* opts, lsg_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* LSG_H_ */

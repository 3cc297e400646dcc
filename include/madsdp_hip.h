/*
 * madsdp_hip.h -- C ABI of the MI355X (gfx950) implementation of MadSDP's
 * interior-point hot path.
 *
 * Two layers:
 *
 *  1. Solver layer (msdp_solver_*): one opaque solver owns a problem
 *     (LowRankOpt.Model fields) and its iterate, resident in device memory,
 *     and exposes the algorithmic-component seam of MadSDP: the driver steps
 *     of src/solver.jl and the component methods of src/components.jl.
 *     A Julia component binds these with `ccall`.
 *
 *  2. Kernel layer (msdp_k_*): the device building blocks on caller-owned
 *     device pointers, on a caller-given HIP stream -- the replacements of
 *     the reference's LAPACK/BLAS calls, exported for unit tests and for
 *     callers that keep their own state.
 *
 * Conventions: matrices are column-major FP64 with explicit leading
 * dimension (Julia's Matrix{Float64}).  Return value 0 = success; < 0 =
 * API/HIP error (msdp_last_error() has the message); > 0 = numerical failure
 * (the reference's `false`, e.g. the 1-based first non-positive pivot).
 * Status codes are MadNLP.Status values (src/moi.jl:16-27).
 */
#ifndef MADSDP_HIP_H
#define MADSDP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msdp_solver msdp_solver;

/* component selectors (src/components.jl) */
enum {
  MSDP_ALG_PRIMAL_DUAL = 0,     /* PrimalDual      components.jl:30-36 */
  MSDP_ALG_DUAL_CENTERING = 1,  /* DualCentering   components.jl:57-65 */
  MSDP_DIR_NT = 0,              /* NT              components.jl:96 */
  MSDP_DIR_HKM = 1,             /* HKM             components.jl:110 */
  MSDP_STEP_EIGEN = 0,          /* EigenStep       components.jl:132-135 */
  MSDP_STEP_BACKTRACK = 1,      /* CholeskyBacktrack components.jl:146-150 */
  MSDP_CENTER_MEHROTRA = 0,     /* Mehrotra        components.jl:169-172 */
  MSDP_CENTER_ADAPTIVE = 1,     /* AdaptiveMu      components.jl:188-192 */
  MSDP_INIT_SCALED_IDENTITY = 0,/* ScaledIdentity  components.jl:208-217 */
  MSDP_INIT_POINT = 1,          /* InitialPoint    components.jl:243-251 */
  MSDP_INIT_DUAL_SHIFT = 2,     /* DualShift       components.jl:228-232 */
  MSDP_SCHUR_DENSE_CHOLESKY = 0,/* DenseCholesky   components.jl:273 */
  MSDP_SCHUR_QR = 1,            /* QRSchur         components.jl:300 */
  MSDP_SCHUR_LDLT = 2           /* MadNLPSchur     components.jl:275-283, kkt.jl:7-32:
                                   Bunch-Kaufman LDL' on the device */
};

/* MadSDPOptions (src/options.jl:3-15) and the component parameters. */
typedef struct msdp_options {
  double tol;              /* 1e-7 */
  int32_t max_iter;        /* 100 */
  double max_wall_time;    /* 3600 */
  double mu_min;           /* 1e-14 */
  double regularize_schur; /* 1e-12 */
  int32_t max_schur_reg;   /* 8 */
  int32_t algorithm, direction, step, centering, initial_point, schur;
  double tau;              /* EigenStep tau 0.95 */
  double shrink, boundary; /* CholeskyBacktrack 0.8, 0.95 */
  double sigma_exponent;   /* Mehrotra 3 */
  double initial_ratio, ratio; /* AdaptiveMu 0.25, 0.5 */
  double identity_scale;   /* ScaledIdentity: <= 0 means automatic */
  double initial_shift, max_shift; /* DualShift 1, 1e10 */
  double dc_beta, dc_delta, dc_A;  /* DualCentering 0.2, 1, 2 (dc_A < 0: theoretical) */
  int32_t dc_centering_max_iter;   /* 50 */
  /* precision of the Schur assembly: 0 = FP64 (the reference's arithmetic);
   * 1 = mixed FP32: FP32 MFMA assembly of the dense-constraint blocks while
   * mu >= mu_switch, FP64 Cholesky, and every Schur solve refined against
   * the FP64 operator (at most refine_max_iter steps, FP64 assembly if the
   * refinement stalls); 2 = mixed BF16x3: the same with the FP64 scaled
   * constraint panel B split into bf16 hi + lo and H = B'B formed by three
   * bf16 MFMA products (FP32 accumulation), the refinement applying the
   * FP64 operator through B.  Not in the reference: an extension of this
   * library. */
  int32_t precision;               /* 0 */
  double mu_switch;                /* 1e-5 */
  int32_t refine_max_iter;         /* 10 */
  /* mixed precision only when one Schur assembly of the dense-constraint
   * blocks costs at least this many GFLOP (4 m_i n^3 + m_i^2 n(n+1)/2 summed
   * over the blocks): below it the refinement costs more than the FP32
   * assembly saves and the solve stays FP64; it also stays FP64 when m
   * exceeds the sum of the blocks' n(n+1)/2 and n_s (a singular H) */
  double lowp_min_gflop;           /* 500 */
  /* 1: the solve shares the device with other solves (a serving loop): every
   * product is one launch (no split-K or stream-K partition of a product over
   * the whole chip, which costs extra launches and pool allocations that
   * concurrent solves pay for in dispatch contention).  Results differ from
   * the shared_device = 0 path only in the summation order of those products.
   * Not in the reference: an extension of this library. */
  int32_t shared_device;           /* 0 */
} msdp_options;

/* iteration data of MadSDPSolver (src/solver.jl:37-48) */
typedef struct msdp_info {
  double mu, sigma, alpha, beta, inf_pr, inf_du, inf_compl, obj_val;
  double pobj, dobj;
  int32_t iter, status, n_newton, n_predictor;
  double dc_t, dc_lambda;
  int32_t lowp_factorizations, fp64_fallbacks, refine_steps, pad;
  /* device bytes the solver holds now (its allocations and the recorder's
   * argument buffer; GEMM partials taken from the stream-ordered pool for the
   * duration of one product are not counted) -- compare msdp_solver_footprint */
  int64_t device_bytes;
  /* the diagonal shift the last Schur factorisation took: 0, or the rung of
   * the regularisation ladder (src/solver.jl:299-314) it succeeded at */
  double schur_reg;
} msdp_info;

/* which iterate array msdp_solver_get/set moves */
enum {
  MSDP_X = 0, MSDP_S = 1, MSDP_Y = 2, MSDP_XLIN = 3, MSDP_ZLIN = 4,
  MSDP_DELX = 5, MSDP_DELS = 6, MSDP_DELY = 7, MSDP_DELXLIN = 8, MSDP_DELZLIN = 9,
  MSDP_RD = 10, MSDP_RP = 11, MSDP_RDLIN = 12, MSDP_W = 13, MSDP_G = 14, MSDP_SI = 15,
  MSDP_CORR = 16, MSDP_LS = 17, MSDP_LX = 18, MSDP_H = 19, MSDP_D = 20
};

const char* msdp_last_error(void);
void msdp_default_options(msdp_options* opt);

/* ---- problem: LRO.Model(C, A, b, d_lin, C_lin, dims) (test/runtests.jl:16-31) ---- */
/* replaces MadSDPSolver(model; ...) construction, src/solver.jl:87-160 */
int msdp_solver_create(msdp_solver** out, int device, int32_t nblocks, const int64_t* dims,
                       int64_t m, int64_t n_scalar);
int msdp_solver_destroy(msdp_solver* s);
/* LRO.grad(model, MatrixIndex(i)): C_i, n_i x n_i */
int msdp_solver_set_cost_block(msdp_solver* s, int32_t block, const double* C);
/* A_ij as COO entries of both triangles (LRO.jac sparse); the one-hot diagonal
 * case (max-cut) is detected and given the Hadamard Schur kernel */
int msdp_solver_set_constraints_coo(msdp_solver* s, int32_t block, int64_t nnz,
                                    const int64_t* con, const int64_t* row, const int64_t* col,
                                    const double* val);
/* A_ij dense for the m_i constraints con_idx (sorted) touching the block;
 * mats is m_i consecutive n_i x n_i matrices (BufferedModelForSchur dense path) */
int msdp_solver_set_constraints_dense(msdp_solver* s, int32_t block, int64_t m_i,
                                      const int64_t* con_idx, const double* mats);
/* Low-rank constraint matrices of one block, A_j = sum_r w_r f_r f_r' (the
 * rank-one / factored matrices LowRankOpt keeps, reference README.md:148):
 * con_idx (m_i, strictly increasing), rank_ptr (m_i + 1, from 0) the column
 * range of each constraint in factors (n x rank_ptr[m_i], column-major), and
 * weights (rank_ptr[m_i]).  A(X), A*(y) and the Schur contribution are then
 * computed from F' X F, F diag(w y) F' and F' L F, F' R F. */
int msdp_solver_set_constraints_lowrank(msdp_solver* s, int32_t block, int64_t m_i,
                                        const int64_t* con_idx, const int64_t* rank_ptr,
                                        const double* factors, const double* weights);
/* LRO.cons_constant (b), d_lin, C_lin (m x n_s column-major) */
int msdp_solver_set_rhs(msdp_solver* s, const double* b);
int msdp_solver_set_scalar_data(msdp_solver* s, const double* d_lin, const double* C_lin);
int msdp_solver_configure(msdp_solver* s, const msdp_options* opt);
/* InitialPoint(X, y; x): X blocks concatenated, y (m), x (n_s) or NULL */
int msdp_solver_set_initial_point(msdp_solver* s, const double* X, const double* y,
                                  const double* x);

/* ---- driver (src/solver.jl:286-396) ---- */
int msdp_solver_initialize(msdp_solver* s);          /* initialize!        solver.jl:286 */
int msdp_solver_iterate(msdp_solver* s);             /* iterate!           solver.jl:316 */
int msdp_solver_update_termination(msdp_solver* s);  /* update_termination! solver.jl:349 */
/* solver.mu = mu (the field MadSDPSolver holds, src/solver.jl:37-48; the
 * reference's test/termination.jl:22 assigns it before update_termination!) */
int msdp_solver_set_mu(msdp_solver* s, double mu);
int msdp_solver_solve(msdp_solver* s);               /* solve!             solver.jl:385 */
int msdp_solver_get_info(const msdp_solver* s, msdp_info* out);
/* number of iteration records and the per-iteration trace (pobj, dobj,
 * inf_pr, inf_du, gap, mu, alpha, beta, sigma) */
int msdp_solver_trace(const msdp_solver* s, int32_t maxrec, double* out9, int32_t* nrec);

/* ---- component seam (src/primal_dual.jl, src/step.jl, src/directions.jl) ---- */
int msdp_solver_update_residuals(msdp_solver* s);                 /* step.jl:147-182 */
int msdp_solver_prepare(msdp_solver* s, int32_t* ok);             /* primal_dual.jl:29-37 */
int msdp_solver_factorize_schur(msdp_solver* s, int32_t* ok);     /* solver.jl:299-314 */
int msdp_solver_solve_directions(msdp_solver* s, double mu, int32_t corrected); /* primal_dual.jl:44-54 */
int msdp_solver_correction(msdp_solver* s);                       /* step.jl:263-271 */
int msdp_solver_step_lengths(msdp_solver* s, double* alpha, double* beta); /* step.jl:186-228 */
int msdp_solver_affine_complementarity(msdp_solver* s, double alpha, double beta,
                                       double* out);              /* step.jl:231-243 */
int msdp_solver_take_step(msdp_solver* s, double alpha, double beta); /* primal_dual.jl:56-72 */
int msdp_solver_schur_solve(msdp_solver* s, double* rhs_host);    /* schur_solvers.jl:21 */
/* copy an iterate array (block index for matrix arrays, ignored otherwise) */
int msdp_solver_get(const msdp_solver* s, int32_t which, int32_t block, double* host);
int msdp_solver_set(msdp_solver* s, int32_t which, int32_t block, const double* host);
/* wait for the solver's stream */
int msdp_solver_synchronize(msdp_solver* s);
/* live timing of one kernel class on the solver's stream (0 = FP64 MFMA GEMM,
 * 1 = tridiagonalisation update, 2 = FP32 MFMA GEMM, 3 = one-workgroup
 * tridiagonalisation): HIP events around every launch, summed
 * duration, launch count and algorithmic work (flops or bytes).  No
 * counterpart in the reference; it feeds the roofline report. */
int msdp_solver_profile_begin(msdp_solver* s, int32_t kernel_class);
int msdp_solver_profile_end(msdp_solver* s, double* total_ms, int64_t* launches, double* work);
/* the same over every stream of the process (the serving loop's concurrent
 * solves); class 3 = the one-workgroup tridiagonalisation (n <= 640).  No
 * counterpart in the reference. */
/* Kernel-selection parameters (no counterpart in the reference): the GEMM
 * tile / split-K / stream-K choices, by name -- gemm_wide_min_tiles,
 * gemm_wide_deep_tiles, splitk_workgroups, splitk_min_k, streamk_occupancy,
 * streamk_min_slabs, streamk_max_waves.  The autotuner's table (madsdp_amd/tuning_gfx950.json)
 * is applied through msdp_set_tuning once, before any solver runs; -2 for an
 * unknown name or a value out of range. */
int msdp_set_tuning(const char* name, double value);
int msdp_get_tuning(const char* name, double* value);
int msdp_profile_begin(int32_t kernel_class);
int msdp_profile_end(double* total_ms, int64_t* launches, double* work);
/* Run the solver on a caller-owned HIP stream (before msdp_solver_initialize;
 * the solver never destroys it).  Without it the solver creates its own
 * non-blocking stream.  No counterpart in the reference: Julia runs one
 * solver per task, the serving loop below runs one per stream. */
int msdp_solver_set_stream(msdp_solver* s, void* stream);
/* Allocate the device state and upload the problem data now (otherwise done
 * by the first msdp_solver_initialize).  No counterpart in the reference. */
int msdp_solver_upload(msdp_solver* s);
/* Device bytes the solver holds once initialized, estimated from the sizes
 * alone (admission by memory footprint; no counterpart in the reference). */
int msdp_solver_footprint(const msdp_solver* s, int64_t* bytes);
/* Copy the solution (X, S, y, x, z) to the host and free the device state;
 * afterwards msdp_solver_get serves those arrays, get_info and trace keep
 * working, and every other call fails.  No counterpart in the reference. */
int msdp_solver_release(msdp_solver* s);

/* ---- serving loop: continuous batching of independent solves ----
 * No counterpart in the reference (MadSDP solves one problem per call,
 * src/solver.jl:385-396).  Requests are prepared solvers (problem data set,
 * options configured, nothing on the device yet) with an arrival time in
 * seconds after msdp_server_start.  `slots` worker threads each own a HIP
 * stream; a free slot takes the first request, in submission order, that has
 * arrived and whose footprint fits in the memory budget next to the running
 * solves (first fit: a request that does not fit is passed over by smaller
 * ones for at most one second, then waits at the head until memory frees; a
 * request larger than the budget runs alone).  An admitted request is uploaded, solved on its
 * slot's stream concurrently with the other slots, and released.  The caller
 * keeps ownership of the solver handles and must not use a submitted one
 * until its request is done. */
typedef struct msdp_server msdp_server;
typedef struct {
  double arrival, admitted, uploaded, done;  /* seconds after msdp_server_start */
  int64_t footprint;               /* bytes */
  int32_t rc;                      /* msdp_solver_solve's return (0 = ok) */
  int32_t slot;
  int32_t finished;                /* 1 once done */
  int32_t pad;
} msdp_request_info;
typedef struct {
  int64_t submitted, completed, peak_bytes;
  int32_t slots, peak_active;
  double elapsed;                  /* seconds since start */
} msdp_server_stats;
int msdp_server_create(msdp_server** out, int device, int32_t slots, int64_t mem_budget_bytes);
int msdp_server_start(msdp_server* v);
int msdp_server_submit(msdp_server* v, msdp_solver* s, double arrival_s, int64_t* id);
/* block until every submitted request is done */
int msdp_server_wait_all(msdp_server* v);
int msdp_server_request(const msdp_server* v, int64_t id, msdp_request_info* out);
int msdp_server_get_stats(const msdp_server* v, msdp_server_stats* out);
int msdp_server_destroy(msdp_server* v);

/* ---- kernel layer: device pointers, caller's stream ---- */
/* C = alpha op(A) op(B) + beta C (BLAS dgemm; directions.jl, step.jl products) */
int msdp_k_gemm(void* stream, int32_t transA, int32_t transB, int32_t M, int32_t N, int32_t K,
                double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                double beta, double* C, int64_t ldc);
/* msdp_k_gemm with triangular operands and an optional lower-triangle output.
 * No counterpart in the reference (BLAS has trmm for one triangular operand):
 * it is the product the solver uses for its triangular factors (L_X V,
 * L_X^-1 dX L_X^-T, L_S^-T L_S^-1, ...).  tri: MSDP_TRI_* bits saying that
 * op(A) / op(B) is lower or upper triangular; every output tile then runs only
 * the K range where both can be nonzero, so the entries on the zero side are
 * never read.  lower = 1: only C's lower triangle is written (and computed). */
enum { MSDP_TRI_A_LOWER = 1, MSDP_TRI_A_UPPER = 2, MSDP_TRI_B_LOWER = 4, MSDP_TRI_B_UPPER = 8 };
int msdp_k_gemm_tri(void* stream, int32_t transA, int32_t transB, int32_t M, int32_t N, int32_t K,
                    double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                    double beta, double* C, int64_t ldc, int32_t lower, int32_t tri);
/* msdp_k_gemm that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial results of the stream-K and split-K paths go to one library-owned
 * device array of msdp_k_gemm_noalloc_scratch_bytes() (128 MiB) per device
 * instead of the stream-ordered pool; a product whose partials exceed it
 * returns -2 with the sizes in msdp_last_error().  Results are bit-identical
 * to msdp_k_gemm's.  Every call on a device shares that device's array, so
 * calls that overlap on one device (other streams, other threads) race. */
int msdp_k_gemm_noalloc(void* stream, int32_t transA, int32_t transB, int32_t M, int32_t N,
                        int32_t K, double alpha, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double beta, double* C, int64_t ldc);
size_t msdp_k_gemm_noalloc_scratch_bytes(void);
/* lower triangle of C = alpha op(A)' op(A) + beta C, upper triangle untouched
 * (BLAS dsyrk; the Schur assembly H = B'B behind LRO.schur_complement!,
 * src/solver.jl:304); trans = 1: C = A' A with A K x N; trans = 0: C = A A' with A N x K */
int msdp_k_syrk_lower(void* stream, int32_t trans, int32_t N, int32_t K, double alpha,
                      const double* A, int64_t lda, double beta, double* C, int64_t ldc);
/* lower Cholesky in place; *info_dev = 0 or first failing pivot (LAPACK potrf) */
int msdp_k_potrf(void* stream, int32_t n, double* A, int64_t lda, int32_t* info_dev,
                 double* work);
size_t msdp_k_potrf_work(int32_t n);
/* _cholesky_factor! (blocks.jl:94-113): L = chol((M+M')/2) + flush; returns the
 * info through *info_dev */
int msdp_k_cholesky_factor(void* stream, int32_t n, const double* M, int64_t ldm, double* L,
                           int64_t ldl, int32_t* info_dev, double* scratch);
size_t msdp_k_cholesky_factor_scratch(int32_t n);
/* Linv = L^-1 for lower-triangular L (LAPACK trtri) */
int msdp_k_trtri(void* stream, int32_t n, const double* L, int64_t ldl, double* Linv,
                 int64_t ldi, double* work);
size_t msdp_k_trtri_work(int32_t n);
/* symmetric indefinite LDL' with Bunch-Kaufman pivoting of the lower triangle
 * in place (LAPACK sytrf 'L', behind MadNLPSchur's BUNCHKAUFMAN solver,
 * src/kkt.jl:15-25): ipiv as LAPACK's (1-based, negative for 2x2 pivots),
 * work 2n doubles, ctl msdp_k_sytrf_ctl_bytes() of device memory, *info_dev =
 * first exactly zero pivot (1-based) or 0 */
int msdp_k_sytrf(void* stream, int32_t n, double* A, int64_t lda, int32_t* ipiv, void* ctl,
                 double* work, int32_t* info_dev);
size_t msdp_k_sytrf_ctl_bytes(void);
/* b = A^-1 b with that factorisation (LAPACK sytrs 'L') */
int msdp_k_sytrs(void* stream, int32_t n, const double* A, int64_t lda, const int32_t* ipiv,
                 double* b);
/* eigenvalues ascending (+ vectors) of a symmetric matrix (LAPACK syevd / LRO.eigen!) */
int msdp_k_syevd(void* stream, int32_t n, const double* A, int64_t lda, double* w, double* V,
                 int64_t ldv, int32_t vectors, double* work, int32_t* iwork);
size_t msdp_k_syevd_work(int32_t n);
size_t msdp_k_syevd_iwork(int32_t n);
/* smallest eigenvalue of a symmetric matrix into *out_dev (LRO.lanczos_eigmin!) */
int msdp_k_eig_min(void* stream, int32_t n, const double* A, int64_t lda, double* out_dev,
                   double* work);
/* the primal and dual step eigenvalues of EigenStep (src/step.jl:180-184: lambda_min
 * of L_X^-1 dX L_X^-T and of L_S^-1 dS L_S^-T) in the same launches: lambda_min(A0)
 * into *out0_dev, lambda_min(A1) into *out1_dev, each bit-identical to msdp_k_eig_min;
 * work holds msdp_k_syevd_work(n) doubles */
int msdp_k_eig_min2(void* stream, int32_t n, const double* A0, const double* A1, int64_t lda,
                    double* out0_dev, double* out1_dev, double* work);
/* C(FP64) = alpha op(A) op(B) + beta C with FP32 A, B on FP32 MFMA (the
 * mixed-precision Schur assembly product); no counterpart in the reference */
int msdp_k_gemm32(void* stream, int32_t transA, int32_t transB, int32_t M, int32_t N, int32_t K,
                  float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                  double beta, double* C, int64_t ldc);
/* lower triangle of C(FP64) = alpha A'A + beta C by BF16x3 split products
 * (A is K x N FP64: hi = bf16(A), lo = bf16(A - hi), C = alpha (hi'hi + hi'lo
 * + lo'hi) on the bf16 MFMA with FP32 accumulation; the mixed-precision
 * Schur assembly H = B'B behind LRO.schur_complement!, src/solver.jl:304);
 * work: msdp_k_syrk_bf16x3_work_bytes(N, K) bytes of device memory for the
 * split operand; no counterpart in the reference */
int msdp_k_syrk_bf16x3(void* stream, int32_t N, int32_t K, double alpha, const double* A,
                       int64_t lda, double beta, double* C, int64_t ldc, void* work);
size_t msdp_k_syrk_bf16x3_work_bytes(int32_t N, int32_t K);
/* out = B'(B v) (+ out when accumulate) in FP64 for B rows x cols: the FP64
 * Schur operator a BF16x3-assembled solve is refined against; work holds
 * msdp_k_panel_apply_work(rows, cols) doubles; no counterpart in the reference */
int msdp_k_panel_apply(void* stream, int32_t rows, int32_t cols, const double* B, int64_t ldb,
                       const double* v, double* out, int32_t accumulate, double* work);
size_t msdp_k_panel_apply_work(int32_t rows, int32_t cols);
/* Hadamard Schur complement of one-hot constraints: H[i,j] += L[i,j] R[j,i]
 * for i >= j (max-cut H = W o W, src/components.jl:102-104) */
int msdp_k_schur_onehot(void* stream, int32_t m, const double* L, int64_t ldl, const double* R,
                        int64_t ldr, double* H, int64_t ldh);

#ifdef __cplusplus
}
#endif
#endif

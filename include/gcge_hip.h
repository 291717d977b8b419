/*
 * gcge_hip.h -- C ABI of the MI355X (gfx950) backend of the GCG eigensolver.
 *
 * Plain pointers and sizes only; no HIP or torch types cross this
 * boundary.  Multivector and matrix handles are opaque; `coef`, `beta`,
 * `inner_prod` and `qAp` arguments of the table slots are host arrays,
 * exactly as in the reference table (src/ops.h:43-152).
 *
 * Errors: the slots return void, as in the reference, which aborts on a
 * failed assert.  This backend prints the failing call and aborts the
 * same way; gcge_hip_last_error() additionally reports the last recorded
 * failure for callers that install their own abort handler.
 */
#ifndef GCGE_HIP_H
#define GCGE_HIP_H

#include <stdint.h>
#include "gcge_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* The drop-in: a HIP "app" that fills the operation table.            */
/* ------------------------------------------------------------------ */

/* Fills every multivector, vector and matrix slot of `ops` with HIP
 * implementations, and installs the host small-dense table as
 * ops->lapack_ops.  Replaces OPS_CCS_Set (app/app_ccs.c:213-249): like
 * it, it expects ops->lapack_ops == NULL and is followed by OPS_Setup.
 * MatAxpby stays NULL, so the driver never shifts A in place
 * (src/ops_eig_sol_gcg.c:600-630).  MultiVecQtAP is filled with a HIP
 * implementation of the semantics of DefaultMultiVecQtAP
 * (src/ops_multi_vec.c:351-411), including op(A)*P left in mv_ws. */
void OPS_HIP_Set(OPS *ops);
/* Frees the host dense table OPS_HIP_Set installed as ops->lapack_ops, for
 * callers that own the table struct itself (the reference's OPS_Destroy,
 * src/ops.c:140-147, frees it recursively instead). */
void OPS_HIP_Clear(OPS *ops);

/* Host small-dense table over the LAPACKVEC layout (gcge_dense_mv);
 * replaces OPS_LAPACK_Set (app/app_lapack.c:970-1014). */
void OPS_HOSTDENSE_Set(OPS *ops);

/* DenseMatQtAP and DenseMatOrth of the host table (app/app_lapack.c:
 * 24-227 and 653-699). */
void gcge_host_DenseMatQtAP(char ntluA, char nsdC, int nrowsA, int ncolsA,
                            int nrowsC, int ncolsC, double alpha, double *matQ,
                            int ldQ, double *matA, int ldA, double *matP, int ldP,
                            double beta, double *matC, int ldC, double *dbl_ws);
void gcge_host_DenseMatOrth(double *mat, int nrows, int ldm, int start, int *end,
                            double orth_zero_tol, double *dbl_ws, int length,
                            int *int_ws);
void gcge_host_Printf(const char *fmt, ...);
/* Sends the table's Printf output (the driver's monitor lines) to stderr
 * instead of stdout when to_stderr != 0. */
void gcge_host_set_print_target(int to_stderr);
double gcge_host_GetWtime(void);
int gcge_host_GetOptionFromCommandLine(const char *name, char type, void *value,
                                       int argc, char *argv[], OPS *ops);

/* ------------------------------------------------------------------ */
/* Device objects                                                      */
/* ------------------------------------------------------------------ */

/* Uploads a CCS matrix (CCSMAT layout, app/app_ccs.h:20-24) to the
 * current device and returns the handle the table slots take as `mat`.
 * The matrix must be symmetric, as the generalized symmetric-definite
 * problem requires: the device kernels read the compressed columns as
 * compressed rows (the reading the reference's MKL path makes,
 * app/app_ccs.c:78-82).  A single-precision copy of the values is kept
 * when flags & GCGE_HIP_MAT_FP32.  Returns NULL on invalid input. */
#define GCGE_HIP_MAT_FP32 1
/* With flags & GCGE_HIP_MAT_GENERAL the matrix is always multiplied as a
 * general sparse matrix.  Without it, a constant-coefficient 7-point
 * operator on an n^3 grid (n <= 128, point x + n y + n^2 z, one diagonal
 * value and one neighbour value, as the stencil problems' A and B) is
 * recognised and multiplied matrix-free, with the same results bit for
 * bit. */
#define GCGE_HIP_MAT_GENERAL 2
void *HIPCCS_Create(const gcge_ccs_mat *host, int flags);
void HIPCCS_Destroy(void **mat);
/* Number of rows, stored nonzeros (including SELL padding) and device
 * bytes of a matrix handle. */
void HIPCCS_Info(void *mat, int64_t *nrows, int64_t *nnz, int64_t *padded,
                 int64_t *bytes);
/* The grid side n when the matrix is multiplied matrix-free as a
 * constant-coefficient 7-point operator (see GCGE_HIP_MAT_GENERAL), else 0. */
int HIPCCS_Stencil(void *mat);

/* Grid operators assembled on the device, without a host CCS.  The
 * reference builds its matrices on the host and hands over a CCSMAT
 * (app/app_ccs.h:20-24); these give the handle HIPCCS_Create would return
 * for the same operator, bit for bit (HIPCCS_Equal), from the grid size
 * and, for variable coefficients, the N node values only.
 *
 * Interior grid of n points per side, N = n^dim, Dirichlet boundary,
 * point index x + n y + n^2 z, rows in ascending column order.
 *   LAPLACE   (dim 2, 5 points; dim 3, 7 points): 2 dim I - Adj.
 *   MASS      (same stencils): (4 dim I + Adj) / (6 dim).
 *   DIFFUSION dim 2, 5 points: edge weight w_ij = 2 k_i k_j / (k_i + k_j)
 *             from the node coefficients k = coef; A_ij = -w_ij, A_ii = the
 *             weights at i (in the order +x, -x, +y, -y) plus k_i per
 *             missing neighbour.
 *             dim 3, 27 points: the neighbour at offset d couples with
 *             g(d) 2 k_i k_j / (k_i + k_j), g(d) = 1 / |d|^2; A_ii sums, in
 *             ascending offset order, the weights and g(d) k_i for each
 *             missing neighbour, then adds delta_rel times the largest
 *             such diagonal.
 *   DIAGONAL  (any dim, points ignored): diag(coef).
 * coef is a host array of N doubles (DIFFUSION, DIAGONAL), else unused.
 * A single-precision copy is kept when flags & GCGE_HIP_MAT_FP32; the 3D
 * LAPLACE and MASS operators are multiplied matrix-free unless flags &
 * GCGE_HIP_MAT_GENERAL (as for HIPCCS_Create).
 * Returns NULL on an unsupported combination or N >= 2^31. */
#define GCGE_GRID_LAPLACE 0
#define GCGE_GRID_MASS 1
#define GCGE_GRID_DIFFUSION 2
#define GCGE_GRID_DIAGONAL 3
typedef struct gcge_grid_op {
    int kind;
    int dim;
    int points;
    int n;
    const double *coef;
    double delta_rel;
} gcge_grid_op;
void *HIPCCS_CreateGrid(const gcge_grid_op *op, int flags);
/* 1 when two matrix handles hold the same device arrays (slice layout,
 * columns, values and the FP32 copy if both have one), 0 otherwise. */
int HIPCCS_Equal(void *a, void *b);

/* Host <-> device copies of columns [start, end) of a multivector; the
 * host array is column-major with leading dimension ld. */
void HIP_MultiVecSetFromHost(void **mv, int start, int end, const double *src,
                             int64_t ld);
void HIP_MultiVecGetToHost(void **mv, int start, int end, double *dst, int64_t ld);
/* Rows, columns and leading dimension of a multivector handle. */
void HIP_MultiVecInfo(void **mv, int64_t *nrows, int64_t *ncols, int64_t *ldd);

/* One Cholesky QR pass over columns [start, end) of x in the B inner
 * product (B NULL: the Euclidean one), all on the device: G = X^T B X,
 * D = diag(G)^{1/2}, D^{-1} G D^{-1} = R^T R, X <- X D^{-1} R^{-1};
 * mv_ws (at least end - start columns, same rows) is overwritten.  The
 * block-orthonormalization step the driver's GCGE_ORTH_CHOLQR runs twice
 * per block in place of the column-by-column self-orthonormalization of
 * the reference's OrthSelf (src/ops_orth.c:271-368).  Returns 1 when
 * done; 0, with x unchanged, when a column's B-norm is below zero_tol or a
 * pivot of the scaled factorization is not above rel_tol (times the
 * largest scaled diagonal entry); -1 when the block is empty or wider
 * than 1024 columns (nothing done). */
int HIP_MultiVecCholQR(void **x, int start, int end, void *B, void **mv_ws, double zero_tol,
                       double rel_tol);

/* Allocation-free inner products.  No counterpart in the reference: this
 * is for callers that cannot allocate device memory (code that must not
 * call hipMalloc, such as tight serving loops or harnesses that forbid
 * it).  The same semantics as the table's MultiVecInnerProd slot
 * (app/app_lapack.c:299-321: nsdIP 'N' full block, 'S' symmetric, 'D'
 * diagonal; inner_prod a host array with leading dimension ldIP), with
 * the partial sums and the device result kept in a fixed 32 MiB array
 * that each GPU holds until the process ends, instead of buffers the
 * library allocates.  Results are the slot's bit for bit whenever the
 * slot's launch plan fits the array; otherwise the rows are grouped in as
 * many chunks as fit (a different summation grouping).  Returns 0, or -1 when the block or
 * its partial sums do not fit the array, or the arguments are invalid
 * (gcge_hip_last_error says which).  Waits for its stream before it
 * returns; calls that overlap in time on one GPU race on the array. */
int HIP_MultiVecInnerProdNoAlloc(char nsdIP, void **x, void **y, int is_vec, int *start,
                                 int *end, double *inner_prod, int ldIP);

/* Random fill of MultiVecSetRandomValue:
 *   GCGE_HIP_RAND_LIBC   (default) host libc rand()/(RAND_MAX+1), column by
 *                        column, uploaded -- the reference's sequence
 *                        (app/app_lapack.c:322-333), so a run started
 *                        after srand(s) sees the same initial block;
 *   GCGE_HIP_RAND_DEVICE U[0,1) from a counter-based generator on the
 *                        device (SplitMix64 of seed and element index).
 * The setting and the generator's position belong to the calling thread;
 * each call restarts the sequence. */
#define GCGE_HIP_RAND_LIBC 0
#define GCGE_HIP_RAND_DEVICE 1
void gcge_hip_set_random_mode(int mode, uint64_t seed);

/* Each host thread that calls into the library gets its own context: a
 * device stream, staging buffers and a rocBLAS handle.  Tables driven
 * from different threads therefore run concurrently on the device (a
 * multivector or matrix may be used by any thread, one at a time).
 *
 * Waits for all work the calling thread queued. */
void gcge_hip_synchronize(void);
/* Releases the calling thread's context (stream, buffers, handle); the
 * next call from the thread creates a new one.  Worker threads call it
 * before they exit. */
void gcge_hip_thread_release(void);
/* Memory pool of the calling thread (off by default): with on != 0,
 * device and pinned host buffers the library allocates on this thread are
 * cached when freed and reused by the thread's later allocations of about
 * the same size (at least the size asked, at most about twice it) instead
 * of going back to HIP, so a thread that serves request after request stops
 * calling hipMalloc and hipFree (which waits for the device) once its pool
 * holds what a request needs.  At most 16 GiB stays cached per thread; off
 * (or gcge_hip_thread_release) hands the cached blocks back.  The solve
 * service's workers run with it on.  Stats: bytes the thread's pool has
 * allocated from HIP (in use or cached), and bytes cached. */
void gcge_hip_set_memory_pool(int on);
void gcge_hip_memory_pool_stats(int64_t *allocated, int64_t *cached);
/* Copies the last error message (empty if none) into buf; returns its
 * code (0 = none). */
int gcge_hip_last_error(char *buf, int len);
/* Number of the visible device the library uses. */
int gcge_hip_device(void);

/* Kernel timing of the calling thread's launches: when enabled, HIP
 * events are recorded on its stream around each launch of the kernel
 * classes below, and read back
 * with their algorithmic bytes and flops (bytes: every operand read once
 * and every result written once).
 *   kind 0: FP64 SpMM (MatDotMultiVec), 1: FP32 shifted SpMM of the
 *   device BlockPCG, 2: Gram (inner products, both kernels),
 *   3: LinearComb (matrix-core kernel), 4: BLAS-1, 5: vector updates of
 *   the device BlockPCG (p = r + beta p; x += alpha p, r -= alpha w with
 *   r.r), 6: LinearComb by the library GEMM (from 448 Ki rows). */
void gcge_hip_profile_enable(int on);
void gcge_hip_profile_reset(void);
void gcge_hip_profile_read(int kind, int64_t *launches, double *total_ms, double *bytes,
                           double *flops);
/* The same timing for every thread (the solve service's workers and its
 * batcher): with on != 0 the totals are cleared and contexts created
 * afterwards time their launches; a context's totals are added when it is
 * released (a worker's when the service is destroyed). */
void gcge_hip_profile_all(int on);
void gcge_hip_profile_read_all(int kind, int64_t *launches, double *total_ms, double *bytes,
                               double *flops);

/* Launch-shape tuner (no counterpart in the reference, whose kernels are
 * MKL's).  A process-wide table maps (operation, shape class) to a
 * launch choice: op 1 the SpMM's column tile, workgroup size and row-block
 * order, op 3 the row-block order of the FP32 BlockPCG's gathering kernels
 * (all these give the same bits), op 2 the MFMA Gram's row chunking (the
 * chunking groups the partial sums, so Gram results are deterministic for
 * a given table).  With tuning on, the first call of a class missing from
 * the table times every candidate on its own operands and records the
 * fastest; with it off, missing classes use the built-in heuristic.  The
 * table tuned on MI355X (tuning_gfx950.txt beside the library, or the
 * file named by GCGE_TUNE_TABLE; an empty value loads none) is loaded
 * when the first device context is created.  gcge_tune_load merges a file
 * into the table and gcge_tune_save writes it; both return the number of
 * entries or -1.  gcge_tune_entry reads entry i (1 if it exists). */
void gcge_tune_set_mode(int on);
int gcge_tune_get_mode(void);
int gcge_tune_load(const char *path);
int gcge_tune_save(const char *path);
void gcge_tune_clear(void);
int gcge_tune_set(int op, int a, int b, int c, int choice);
int gcge_tune_entry(int i, int *op, int *a, int *b, int *c, int *choice, double *best_us,
                    double *default_us);

/* ------------------------------------------------------------------ */
/* The GCG driver                                                      */
/* ------------------------------------------------------------------ */
/* A reentrant restatement of the reference driver
 * (src/ops_eig_sol_gcg.c:64-1562) that calls only through the table; it
 * runs over OPS_HIP_Set (the FP64 GPU path) or OPS_HOSTDENSE_Set.  The
 * names mirror the reference entry points with the HIPGCG_ prefix so both
 * libraries can be linked into one program. */

/* EigenSolverSetup_GCG (src/ops_eig_sol_gcg.c:1565-1628): stores the
 * sizes, tolerances and workspaces and sets ops->EigenSolver.  mv_ws[0]
 * has nevMax + 2 block_size columns, mv_ws[1..3] block_size columns
 * (as created by HIPGCG_Create). */
void HIPGCG_EigenSolverSetup(double gapMin, int nevInit, int nevMax, int block_size,
                             double tol[2], int numIterMax, int print, int printtime,
                             int user_defined_multi_linear_solver, void **mv_ws[4],
                             double *dbl_ws, int *int_ws, OPS *ops);
/* GCGE_Setparameters (src/ops_eig_sol_gcg.c:1878-1918). */
void HIPGCG_Setparameters(double gapMin, int shift, OPS *ops);
/* EigenSolverSetParameters_GCG (src/ops_eig_sol_gcg.c:1684-1741). */
void HIPGCG_EigenSolverSetParameters(
    int check_conv_max_num, const char *initX_orth_method, int initX_orth_block_size,
    int initX_orth_max_reorth, double initX_orth_zero_tol, const char *compP_orth_method,
    int compP_orth_block_size, int compP_orth_max_reorth, double compP_orth_zero_tol,
    const char *compW_orth_method, int compW_orth_block_size, int compW_orth_max_reorth,
    double compW_orth_zero_tol, int compW_cg_max_iter, double compW_cg_rate,
    double compW_cg_tol, const char *compW_cg_tol_type, int compW_cg_auto_shift,
    int compRR_min_num, double compRR_min_gap, double compRR_tol, OPS *ops);
/* GCGE_Create / GCGE_Destroymvws (src/ops_eig_sol_gcg.c:1865-1877,
 * 1919-1927); dbl_ws and int_ws are sized and allocated here. */
void HIPGCG_Create(void *A, int nevMax, int block_size, int nevInit, void ***mv_ws,
                   double **dbl_ws, int **int_ws, OPS *ops);
void HIPGCG_Destroy(void ***mv_ws, double *dbl_ws, int *int_ws, int nevMax,
                    int block_size, OPS *ops);
/* Releases the solver state stored in ops->eigen_solver_workspace and the
 * device BlockPCG workspace when one is installed. */
void HIPGCG_Free(OPS *ops);
/* Results of the last solve: stats[0] numIter, stats[1] nevConv,
 * stats[2] total BlockPCG iterations (of the driver's FP64 BlockPCG and of
 * a device FP32 BlockPCG in the MultiLinearSolver slot), stats[3] linear
 * solves;
 * times[0..9] phase seconds (initX, checkconv, compP, compRR, rr_matW,
 * dsyevx, compRV, compW, compX, linsol; the names of
 * src/ops_eig_sol_gcg.c:27-41). */
void HIPGCG_GetStats(OPS *ops, int64_t *stats, double *times);

/* Dense symmetric eigensolver used by the Rayleigh-Ritz step:
 * GCGE_EIG_HOST (gcge_host_dense_eigh below) or GCGE_EIG_DEVICE
 * (rocSOLVER dsyevd on the device).  Without a call, the device solver
 * is used for projected matrices of order 128 and above. */
#define GCGE_EIG_HOST 0
#define GCGE_EIG_DEVICE 1
void HIPGCG_SetDenseEigensolver(OPS *ops, int which);
/* The host dense symmetric eigensolver (Householder tridiagonalization
 * and implicit QR): all eigenvalues of the n x n matrix whose upper
 * triangle A holds (lda), ascending in w, orthonormal eigenvectors in Z
 * (ldz).  Returns 0, or 1 if the QR iteration did not converge.  Stands
 * in for the reference's dsyevx call (src/ops_eig_sol_gcg.c:1207). */
int gcge_host_dense_eigh(int n, const double *A, int lda, double *w, double *Z, int ldz);

/* B-orthonormalization of the W block and of the initial block:
 * GCGE_ORTH_MGS (default) is the reference's block modified Gram-Schmidt
 * with column-by-column self-orthonormalization (src/ops_orth.c:203-394);
 * GCGE_ORTH_CHOLQR keeps its removal of the earlier columns but
 * orthonormalizes the block itself by Cholesky QR applied twice, falling
 * back to MGS when the block is numerically rank deficient.  Same
 * subspace, different rounding: results agree in eigenvalues and
 * residuals, not bit for bit.  With the HIP table, GCGE_ORTH_CHOLQR keeps
 * the Gram matrix, its Cholesky factorization and the triangular inverse
 * on the device (blocks of up to 1024 columns; wider ones, and other
 * tables, factor on the host); GCGE_ORTH_CHOLQR_HOST always factors on
 * the host. */
#define GCGE_ORTH_MGS 0
#define GCGE_ORTH_CHOLQR 1
#define GCGE_ORTH_CHOLQR_HOST 2
void HIPGCG_SetOrthMethod(OPS *ops, int which);
/* Counts of the last solve: blocks orthonormalized with GCGE_ORTH_CHOLQR,
 * and how many of them fell back to MGS. */
void HIPGCG_GetOrthStats(OPS *ops, int64_t *cholqr_blocks, int64_t *mgs_fallbacks);

/* ------------------------------------------------------------------ */
/* Mixed precision: FP32 BlockPCG in the MultiLinearSolver slot        */
/* ------------------------------------------------------------------ */
/* Installs a device BlockPCG that runs in single precision as
 * ops->MultiLinearSolver (the user-defined solver of
 * src/ops_eig_sol_gcg.c:585-642, user_defined_multi_linear_solver = 1).
 * A and B must carry FP32 copies (GCGE_HIP_MAT_FP32).  Same stopping
 * rule as BlockPCG (src/ops_lin_sol.c:383-394): per column, residual
 * below rate * initial or below tol, at most max_iter iterations.
 * sigma is the fixed shift of A + sigma B (0 when autoshift is off). */
void HIPGCG_MultiLinearSolverSetup_BlockPCG32(int max_iter, double rate, double tol,
                                              void *B, double sigma, OPS *ops);
/* Frees the device BlockPCG's workspace (four FP32 blocks of the matrix
 * order times the block width) and clears the slot; no counterpart in the
 * reference, whose BlockPCG workspace is a static (src/ops_lin_sol.c:
 * 439-465). */
void HIPGCG_MultiLinearSolverFree_BlockPCG32(OPS *ops);
/* Inner iterations between two reads of the active-column count run as one
 * captured HIP graph (default on; off while kernel timing is enabled).  No
 * counterpart in the reference; for callers that need plain stream launches. */
void HIPGCG_BlockPCG32_SetGraphs(OPS *ops, int on);
/* Up to 65,536 rows an inner iteration is three launches, the two
 * single-workgroup reductions being done by every workgroup of the kernel
 * that follows each (default on; same sums, so the same bits as the
 * five-launch iteration).  No counterpart in the reference. */
void HIPGCG_BlockPCG32_SetFused(OPS *ops, int on);

/* ------------------------------------------------------------------ */
/* Solve service: concurrent requests on one GPU                       */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one problem per process there).  A
 * pool of worker threads, each on its own device stream, takes requests
 * in arrival order as soon as a worker is free and the request's device
 * footprint (gcge_request_footprint) fits in what the HBM budget has
 * left; each request runs this library's driver over its own table. */
#define GCGE_MODE_FP64 0
#define GCGE_MODE_MIXED 1
typedef struct {
    const gcge_ccs_mat *A, *B; /* host CCS; valid until the request completes */
    int nev, block_size, nev_init, nev_max, max_iter;
    int mode;  /* GCGE_MODE_FP64 (autoshift per `shift`, MGS) or
                  GCGE_MODE_MIXED (FP32 BlockPCG, CholQR2, no autoshift) */
    int shift;
    double tol_abs, tol_rel, gap_min;
    uint64_t seed; /* initial block from the device generator */
    /* warm start: nev_given (<= nev_init) initial vectors, N x nev_given
     * column-major with leading dimension N, valid until the request
     * completes; they take the place of the first generated columns, as
     * nevGiven does for the reference's EigenSolver (gcge.c:519). */
    const double *x0;
    int nev_given;
} gcge_request;
typedef struct {
    int status;     /* 1 pending, 0 done, -1 matrix rejected, -2 over budget */
    int nev_conv, num_iter;
    int concurrent; /* requests on the device when this one started, itself included */
    double queue_seconds, setup_seconds, solve_seconds, total_seconds;
    double *eval;   /* caller's array of nev_max doubles, or NULL */
    double *evec;   /* caller's N x nev_max column-major array, or NULL */
    int64_t pool_bytes; /* device and pinned bytes the worker's memory pool
                           holds from HIP after this request (in use or
                           cached for the next requests) */
} gcge_result;
/* Device bytes a request holds while it runs.  A request with nev_given
 * outside [0, nev_init], or nev_given > 0 without x0, is refused
 * (status -1). */
int64_t gcge_request_footprint(const gcge_request *req);
void *gcge_service_create(int workers, int64_t hbm_budget_bytes);
/* Queues a request; returns its ticket, or -1 (res->status says why).  The
 * service writes into *res until gcge_service_wait returns. */
int64_t gcge_service_submit(void *svc, const gcge_request *req, gcge_result *res);
/* Blocks until the request is done; returns its status. */
int gcge_service_wait(void *svc, int64_t ticket);
void gcge_service_stats(void *svc, int64_t *completed, int *max_concurrent,
                        int64_t *hbm_in_use);
/* Finishes the queued requests and joins the workers. */
void gcge_service_destroy(void *svc);
/* Batched inner solves (on by default): the FP32 BlockPCG iterations of
 * the mixed-precision requests in progress (systems of at most 65,536 rows
 * with at most 8 entries per row) run in shared launches, one per step for
 * all of them, driven by one service thread; each request gets the
 * results it would get alone, bit for bit.  Applies to requests started
 * after the call.  Stats: batched rounds, request iterations they carried,
 * and the most requests one round carried. */
void gcge_service_set_batching(void *svc, int on);
/* Block orthonormalization of the mixed-precision requests started after
 * the call: GCGE_ORTH_CHOLQR_HOST (the default: Cholesky factorizations on
 * the host, which serve many concurrent small requests faster) or
 * GCGE_ORTH_CHOLQR (on the device).  Returns 0, or -1 for another value.
 * No counterpart in the reference (its solver has one orthogonalization). */
int gcge_service_set_orth(void *svc, int method);
void gcge_service_batch_stats(void *svc, int64_t *rounds, int64_t *solve_iterations,
                              int *max_solves);

/* ------------------------------------------------------------------ */
/* EPS-style entry point                                               */
/* ------------------------------------------------------------------ */
/* One call with the parameter policy and call sequence of the reference's
 * SLEPc plug-in (EPSSetUp_GCGE / EPSSolve_GCGE, gcge.c:364-535): block
 * size, nevInit and nevMax from nev (and ncv), tolerances {1e-1 absolute,
 * 1e-8 relative} with `tol` replacing one of them, gapMin 1e-5, autoshift
 * on, srand(0) before the solve; nconv = min(converged, nev), reason and
 * its as the plug-in reports them. */
#define GCGE_EPS_RELATIVE 0
#define GCGE_EPS_ABSOLUTE 1
#define GCGE_EPS_CONVERGED_TOL 1   /* EPS_CONVERGED_TOL */
#define GCGE_EPS_DIVERGED_ITS (-1) /* EPS_DIVERGED_ITS */
#define GCGE_EPS_ERR_ARG (-10)
#define GCGE_EPS_ERR_NCV (-11)     /* ncv inconsistent with nev (gcge.c:408-427) */
#define GCGE_EPS_ERR_SIZE (-12)    /* nevMax above the matrix order (gcge.c:490-492) */
#define GCGE_EPS_ERR_MATRIX (-13)  /* HIPCCS_Create rejected A or B */
typedef struct {
    int nev;
    int ncv;           /* 0: the plug-in's default */
    double tol;        /* 0: default */
    int tol_type;      /* GCGE_EPS_RELATIVE (the plug-in's default test) or _ABSOLUTE */
    int max_it;        /* 0: unlimited */
    int autoshift;     /* -1: default (on); ignored (off) in mixed mode */
    int mode;          /* GCGE_MODE_FP64 or GCGE_MODE_MIXED */
    int print;         /* monitor lines through the table's Printf */
    int device_random; /* 0: libc rand() after srand(0), as the plug-in */
    uint64_t seed;     /* device generator seed when device_random */
} gcge_eps_options;
/* Block size, nevInit and nevMax the plug-in derives from nev and ncv;
 * returns 0 or GCGE_EPS_ERR_NCV / _ARG.  No device use. */
int gcge_eps_parameters(int nev, int ncv, int *block_size, int *nev_init, int *nev_max);
/* Solves; eigr receives nconv values (array of nev), evec (or NULL) the
 * N x nconv eigenvectors column-major.  Returns 0 or a GCGE_EPS_ERR_*. */
int gcge_eps_solve(const gcge_ccs_mat *A, const gcge_ccs_mat *B, const gcge_eps_options *opt,
                   double *eigr, double *evec, int *nconv, int *its, int *reason);

#ifdef __cplusplus
}
#endif

#endif /* GCGE_HIP_H */

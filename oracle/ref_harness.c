/*
 * ref_harness.c -- drives the reference GCGE library (compiled from the
 * reference sources into oracle/_ref) for the oracle's checks.
 *
 * Test infrastructure only: it is linked against the reference build and
 * is loaded by tests and by bench.py's cpu_baseline leg, never by the
 * product.
 *
 * ref_solve() mirrors the SLEPc plugin's call sequence
 * (gcge.c:442-535): a random evec block, GCGE_Create, the double/int
 * workspaces sized as in gcge.c:500-507, srand(0), EigenSolverSetup_GCG,
 * GCGE_Setparameters(gapMin, shift) and ops->EigenSolver.  The operator
 * table is either the reference CCS app (app/app_ccs.c:213-249) or the
 * reference dense LAPACK app (app/app_lapack.c:970-1014), or a caller-
 * supplied table (used to run the reference driver over another backend).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ops.h"
#include "app_ccs.h"
#include "app_lapack.h"
#include "ops_eig_sol_gcg.h"

static double wall_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Table constructors exported to ctypes. */
OPS *ref_ops_create_ccs(void)
{
    OPS *ops = NULL;
    OPS_Create(&ops);
    OPS_CCS_Set(ops);
    OPS_Setup(ops);
    return ops;
}

OPS *ref_ops_create_lapack(void)
{
    OPS *ops = NULL;
    OPS_Create(&ops);
    OPS_LAPACK_Set(ops);
    OPS_Setup(ops);
    return ops;
}

/* An empty table (all NULL), for a backend to fill before ref_ops_setup. */
OPS *ref_ops_create_empty(void)
{
    OPS *ops = NULL;
    OPS_Create(&ops);
    return ops;
}

void ref_ops_setup(OPS *ops) { OPS_Setup(ops); }

void ref_ops_destroy(OPS *ops) { OPS_Destroy(&ops); }

/* Host objects in the reference apps' layouts. */
CCSMAT *ref_ccs_wrap(int n, int *j_col, int *i_row, double *data)
{
    CCSMAT *m = (CCSMAT *)malloc(sizeof(CCSMAT));
    m->nrows = n;
    m->ncols = n;
    m->j_col = j_col;
    m->i_row = i_row;
    m->data = data;
    return m;
}

LAPACKMAT *ref_dense_wrap(int nrows, int ncols, int ldd, double *data)
{
    LAPACKMAT *m = (LAPACKMAT *)malloc(sizeof(LAPACKMAT));
    m->nrows = nrows;
    m->ncols = ncols;
    m->ldd = ldd;
    m->data = data;
    return m;
}

void ref_free(void *p) { free(p); }

void ref_srand(unsigned int seed) { srand(seed); }

/* Runs one GCG solve through the table `ops` (already OPS_Setup).
 * A, B are handles understood by that table (B may be NULL).
 * nev: requested pairs; eval has nevMax entries; evec_out (nrows x nevMax,
 * column-major) receives the Ritz vectors when non-NULL and the table
 * uses the host dense layout for multivectors (host_vectors != 0).
 * stats[0] = numIter, stats[1] = nevConv; times[0] = solve seconds.
 * Returns 0. */
int ref_solve(OPS *ops, void *A, void *B, int nrows, int nev, int block_size,
              int nevInit, int nevMax, double tol_abs, double tol_rel,
              double gapMin, int shift, int max_iter, int print,
              double *eval, double *evec_out, int host_vectors,
              int *stats, double *times)
{
    void **evec = NULL;
    void **gcg_mv_ws[4];
    double *dbl_ws;
    int *int_ws;
    double tol[2];
    int nevConv = nev, sizeV, length_dbl_ws, length_int_ws;
    double t0, t1;

    memset(eval, 0, (size_t)nevMax * sizeof(double));
    ops->MultiVecCreateByMat(&evec, nevMax, A, ops);
    ops->MultiVecSetRandomValue(evec, 0, nevMax, ops);
    GCGE_Create(A, nevMax, block_size, nevInit, gcg_mv_ws, NULL, NULL, ops);
    sizeV = nevInit + 2 * block_size;
    length_dbl_ws = 2 * sizeV * sizeV + 10 * sizeV + (nevMax + 2 * block_size)
                    + nevMax * block_size;
    length_int_ws = 6 * sizeV + 2 * (block_size + 3);
    dbl_ws = (double *)calloc((size_t)length_dbl_ws, sizeof(double));
    int_ws = (int *)calloc((size_t)length_int_ws, sizeof(int));
    srand(0);
    tol[0] = tol_abs;
    tol[1] = tol_rel;
    EigenSolverSetup_GCG(gapMin, nevInit, nevMax, block_size, tol, max_iter,
                         print, 0, 0, gcg_mv_ws, dbl_ws, int_ws, ops);
    GCGE_Setparameters(gapMin, shift, ops);

    t0 = wall_now();
    ops->EigenSolver(A, B, eval, evec, 0, &nevConv, ops);
    t1 = wall_now();

    stats[0] = ((GCGSolver *)ops->eigen_solver_workspace)->numIter;
    stats[1] = nevConv;
    times[0] = t1 - t0;
    if (evec_out != NULL && host_vectors) {
        LAPACKVEC *ev = (LAPACKVEC *)evec;
        int c;
        for (c = 0; c < nevMax; ++c)
            memcpy(evec_out + (size_t)nrows * c, ev->data + (size_t)ev->ldd * c,
                   (size_t)nrows * sizeof(double));
    }
    ops->MultiVecDestroy(&evec, nevMax, ops);
    /* GCGE_Destroymvws frees dbl_ws and int_ws as well. */
    GCGE_Destroymvws(gcg_mv_ws, dbl_ws, int_ws, nevMax, block_size, ops);
    return 0;
}

/* Per-phase timers of the reference driver (src/ops_eig_sol_gcg.c:27-41),
 * in the order initX, checkconv, compP, compRR, rr_matW, dsyevx, compRV,
 * compW, compX, linsol.  They are accumulated with ops->GetWtime, which
 * without MPI/OpenMP is process CPU time (src/ops_multi_vec.c:45-56). */
extern struct {
    double v[11];
} time_gcg;

void ref_phase_times(double *out)
{
    memcpy(out, time_gcg.v, sizeof(time_gcg.v));
}

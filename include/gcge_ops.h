/*
 * gcge_ops.h -- the GCGE operation table, declared for the HIP backend.
 *
 * This is the plug-in ABI of the GCG eigensolver: every N-length operation
 * the driver performs goes through one of these function pointers.  The
 * layout (field order, types, and the struct tag `OPS_`) is identical to
 * the reference table in src/ops.h:43-152, so a table filled by this
 * library can be handed to the reference driver (src/ops_eig_sol_gcg.c)
 * unchanged, and a table created by the reference's OPS_Create
 * (src/ops.c:26-85) can be filled by OPS_HIP_Set.  The layout is pinned by
 * offsetof checks in the test suite (tests/test_abi.py) against the
 * reference header compiled in the oracle build.
 *
 * If the reference header has already been included (_OPS_H_ defined),
 * its declaration is used and this one is skipped.
 *
 * Conventions the slots follow (src/ops_multi_vec.c:151-411,
 * app/app_lapack.c:299-534):
 *   - start[2]/end[2] are half-open column ranges; index 0 is the source
 *     (x, or the rows of an inner-product block), index 1 the destination
 *     (y, or the columns).
 *   - nsdIP: 'N' full block ip[r + ldIP*c] = x_r . y_c; 'S' square block,
 *     symmetric result (lower triangle computed, mirrored); 'D' diagonal
 *     ip[ldIP*c] = x_c . y_c.
 *   - coef / beta / inner_prod / qAp are HOST pointers; multivector and
 *     matrix handles are opaque.
 */
#ifndef GCGE_OPS_H
#define GCGE_OPS_H

#ifndef _OPS_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OPS_ {
    void   (*Printf)(const char *fmt, ...);
    double (*GetWtime)(void);
    int    (*GetOptionFromCommandLine)(const char *name, char type, void *data,
                                       int argc, char *argv[], struct OPS_ *ops);
    /* matrices */
    void (*MatView)(void *mat, struct OPS_ *ops);
    void (*MatAxpby)(double alpha, void *matX, double beta, void *matY, struct OPS_ *ops);
    /* single vectors */
    void (*VecCreateByMat)(void **des_vec, void *src_mat, struct OPS_ *ops);
    void (*VecCreateByVec)(void **des_vec, void *src_vec, struct OPS_ *ops);
    void (*VecDestroy)(void **des_vec, struct OPS_ *ops);
    void (*VecView)(void *x, struct OPS_ *ops);
    void (*VecInnerProd)(void *x, void *y, double *inner_prod, struct OPS_ *ops);
    void (*VecLocalInnerProd)(void *x, void *y, double *inner_prod, struct OPS_ *ops);
    void (*VecSetRandomValue)(void *x, struct OPS_ *ops);
    void (*VecAxpby)(double alpha, void *x, double beta, void *y, struct OPS_ *ops);
    void (*MatDotVec)(void *mat, void *x, void *y, struct OPS_ *ops);
    void (*MatTransDotVec)(void *mat, void *x, void *y, struct OPS_ *ops);
    /* multivectors */
    void (*MultiVecCreateByMat)(void ***multi_vec, int num_vec, void *src_mat, struct OPS_ *ops);
    void (*MultiVecCreateByVec)(void ***multi_vec, int num_vec, void *src_vec, struct OPS_ *ops);
    void (*MultiVecCreateByMultiVec)(void ***multi_vec, int num_vec, void **src_mv, struct OPS_ *ops);
    void (*MultiVecDestroy)(void ***multi_vec, int num_vec, struct OPS_ *ops);
    void (*GetVecFromMultiVec)(void **multi_vec, int col, void **vec, struct OPS_ *ops);
    void (*RestoreVecForMultiVec)(void **multi_vec, int col, void **vec, struct OPS_ *ops);
    void (*MultiVecView)(void **x, int start, int end, struct OPS_ *ops);
    void (*MultiVecLocalInnerProd)(char nsdIP, void **x, void **y, int is_vec,
                                   int *start, int *end, double *inner_prod, int ldIP,
                                   struct OPS_ *ops);
    void (*MultiVecInnerProd)(char nsdIP, void **x, void **y, int is_vec,
                              int *start, int *end, double *inner_prod, int ldIP,
                              struct OPS_ *ops);
    void (*MultiVecSetRandomValue)(void **multi_vec, int start, int end, struct OPS_ *ops);
    void (*MultiVecAxpby)(double alpha, void **x, double beta, void **y,
                          int *start, int *end, struct OPS_ *ops);
    void (*MultiVecLinearComb)(void **x, void **y, int is_vec, int *start, int *end,
                               double *coef, int ldc, double *beta, int incb,
                               struct OPS_ *ops);
    void (*MatDotMultiVec)(void *mat, void **x, void **y, int *start, int *end,
                           struct OPS_ *ops);
    void (*MatTransDotMultiVec)(void *mat, void **x, void **y, int *start, int *end,
                                struct OPS_ *ops);
    void (*MultiVecQtAP)(char ntsA, char ntsdQAP, void **mvQ, void *matA, void **mvP,
                         int is_vec, int *start, int *end, double *qAp, int ldQAP,
                         void **mv_ws, struct OPS_ *ops);
    /* host small-dense operations */
    struct OPS_ *lapack_ops;
    void (*DenseMatQtAP)(char ntluA, char nsdC, int nrowsA, int ncolsA,
                         int nrowsC, int ncolsC, double alpha,
                         double *matQ, int ldQ, double *matA, int ldA,
                         double *matP, int ldP, double beta, double *matC, int ldC,
                         double *dbl_ws);
    void (*DenseMatOrth)(double *mat, int nrows, int ldm, int start, int *end,
                         double orth_zero_tol, double *dbl_ws, int length, int *int_ws);
    /* linear solvers */
    void (*LinearSolver)(void *mat, void *b, void *x, struct OPS_ *ops);
    void *linear_solver_workspace;
    void (*MultiLinearSolver)(void *mat, void **b, void **x, int *start, int *end,
                              struct OPS_ *ops);
    void *multi_linear_solver_workspace;
    /* orthonormalization */
    void (*MultiVecOrth)(void **x, int start_x, int *end_x, void *B, struct OPS_ *ops);
    void *orth_workspace;
    /* multigrid (not provided by the HIP backend) */
    void (*MultiGridCreate)(void ***A_array, void ***B_array, void ***P_array,
                            int *num_levels, void *A, void *B, struct OPS_ *ops);
    void (*MultiGridDestroy)(void ***A_array, void ***B_array, void ***P_array,
                             int *num_levels, struct OPS_ *ops);
    void (*VecFromItoJ)(void **P_array, int level_i, int level_j, void *vec_i,
                        void *vec_j, void **vec_ws, struct OPS_ *ops);
    void (*MultiVecFromItoJ)(void **P_array, int level_i, int level_j,
                             void **multi_vec_i, void **multi_vec_j, int *startIJ,
                             int *endIJ, void ***multi_vec_ws, struct OPS_ *ops);
    /* eigensolver */
    void (*EigenSolver)(void *A, void *B, double *eval, void **evec, int nevGiven,
                        int *nevConv, struct OPS_ *ops);
    void *eigen_solver_workspace;
    /* nested application table */
    struct OPS_ *app_ops;
} OPS;

#ifdef __cplusplus
}
#endif

#endif /* _OPS_H_ */

/*
 * Host-side data layouts shared with the reference apps.
 *
 * gcge_ccs_mat has the layout of CCSMAT (app/app_ccs.h:20-24): 0-based
 * compressed columns, column j holds rows i_row[j_col[j] .. j_col[j+1]-1].
 * gcge_dense_mv has the layout of LAPACKVEC/LAPACKMAT (app/app_lapack.h:
 * 17-20): column-major, element (r, c) at data[r + ldd*c].  The GCG driver
 * builds such structs itself for its small host matrices
 * (src/ops_eig_sol_gcg.c:386-395) and hands them to ops->lapack_ops, so
 * the host dense table must read exactly this layout.
 */
typedef struct gcge_ccs_mat_ {
    double *data;
    int    *i_row;
    int    *j_col;
    int     nrows;
    int     ncols;
} gcge_ccs_mat;

typedef struct gcge_dense_mv_ {
    double *data;
    int     nrows;
    int     ncols;
    int     ldd;
} gcge_dense_mv;

#endif /* GCGE_OPS_H */

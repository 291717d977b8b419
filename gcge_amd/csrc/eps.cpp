// EPS-style entry point: one call that solves A x = lambda B x with the
// parameter policy and call sequence of the reference's SLEPc plug-in
// (EPSSetUp_GCGE / EPSSolve_GCGE, gcge.c:364-535), over this library's
// table and driver.  For callers that want the plug-in's behaviour without
// PETSc/SLEPc.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gcge_hip.h"

namespace {

// gcge.c:394-431: block size, nevInit and nevMax from nev (and ncv).
// Returns 0 or the plug-in's error for an inconsistent ncv.
int block_policy(int nev, int ncv, int *bs, int *nev_init, int *nev_max)
{
    if (nev > 6) {
        if (nev < 20)
            *bs = nev / 3;
        else if (nev < 1000)
            *bs = nev / 5;
        else if (nev < 4000)
            *bs = 100;
        else
            *bs = 200;  // the plug-in leaves nev >= 10000 unset; 200 continues its pattern
        *nev_init = 3 * *bs;
        *nev_max = *nev_init + nev;
        if (ncv > 0) *nev_max = ncv;
        if (*nev_init > *nev_max) return GCGE_EPS_ERR_NCV;        // ncv >= nevInit
        if (*nev_max > nev + *nev_init) return GCGE_EPS_ERR_NCV;  // ncv <= nevConv + nevInit
    } else {
        *bs = nev;
        *nev_init = 2 * nev;
        *nev_max = *nev_init;
        if (ncv > 0 && ncv != *nev_init) return GCGE_EPS_ERR_NCV;  // ncv == 2 nev
    }
    return 0;
}

}  // namespace

extern "C" {

int gcge_eps_parameters(int nev, int ncv, int *block_size, int *nev_init, int *nev_max)
{
    if (nev < 1) return GCGE_EPS_ERR_ARG;
    return block_policy(nev, ncv, block_size, nev_init, nev_max);
}

int gcge_eps_solve(const gcge_ccs_mat *A, const gcge_ccs_mat *B, const gcge_eps_options *o,
                   double *eigr, double *evec, int *nconv, int *its, int *reason)
{
    if (A == nullptr || o == nullptr || eigr == nullptr || o->nev < 1) return GCGE_EPS_ERR_ARG;
    int bs, nev_init, nev_max;
    const int st = block_policy(o->nev, o->ncv, &bs, &nev_init, &nev_max);
    if (st != 0) return st;
    if (nev_max > A->nrows) return GCGE_EPS_ERR_SIZE;  // gcge.c:490-492
    // tolerances (gcge.c:376-387): {1e-1 absolute, 1e-8 relative}, one of
    // them replaced by tol according to the convergence test
    double tol[2] = {1e-1, 1e-8};
    if (o->tol > 0) tol[o->tol_type == GCGE_EPS_ABSOLUTE ? 0 : 1] = o->tol;
    const int max_it = o->max_it > 0 ? o->max_it : 0x7fffffff;
    const double gap_min = 1e-5;
    const bool mixed = o->mode == GCGE_MODE_MIXED;
    // the reference forbids autoshift with a user solver
    // (src/ops_eig_sol_gcg.c:503): the mixed path runs without it
    const int shift = mixed ? 0 : (o->autoshift >= 0 ? o->autoshift : 1);
    const int flags = mixed ? GCGE_HIP_MAT_FP32 : 0;
    void *dA = HIPCCS_Create(A, flags);
    if (dA == nullptr) return GCGE_EPS_ERR_MATRIX;
    void *dB = nullptr;
    if (B != nullptr) {
        dB = HIPCCS_Create(B, flags);
        if (dB == nullptr) {
            HIPCCS_Destroy(&dA);
            return GCGE_EPS_ERR_MATRIX;
        }
    }
    OPS *ops = static_cast<OPS *>(calloc(1, sizeof(OPS)));
    OPS_HIP_Set(ops);
    std::vector<double> eval((size_t)nev_max, 0.0);
    void **ev = nullptr;
    ops->MultiVecCreateByMat(&ev, nev_max, dA, ops);
    void **mv_ws[4] = {nullptr, nullptr, nullptr, nullptr};
    HIPGCG_Create(dA, nev_max, bs, nev_init, mv_ws, nullptr, nullptr, ops);
    if (o->device_random) {
        gcge_hip_set_random_mode(GCGE_HIP_RAND_DEVICE, o->seed);
    } else {
        gcge_hip_set_random_mode(GCGE_HIP_RAND_LIBC, 0);
        srand(0);  // gcge.c:508
    }
    HIPGCG_EigenSolverSetup(gap_min, nev_init, nev_max, bs, tol, max_it, o->print ? 1 : 0, 0,
                            mixed ? 1 : 0, mv_ws, nullptr, nullptr, ops);
    HIPGCG_Setparameters(gap_min, shift, ops);
    if (mixed) {
        HIPGCG_SetOrthMethod(ops, GCGE_ORTH_CHOLQR);
        HIPGCG_MultiLinearSolverSetup_BlockPCG32(30, 1e-2, 1e-14, dB, 0.0, ops);
    }
    int nc = o->nev;
    ops->EigenSolver(dA, dB, eval.data(), ev, 0, &nc, ops);
    int64_t stats[4];
    double times[10];
    HIPGCG_GetStats(ops, stats, times);
    // gcge.c:525-531
    const int k = nc > o->nev ? o->nev : nc;
    *nconv = k;
    *reason = k >= o->nev ? GCGE_EPS_CONVERGED_TOL : GCGE_EPS_DIVERGED_ITS;
    *its = (int)stats[0] + (k >= o->nev ? 1 : 0);
    memcpy(eigr, eval.data(), sizeof(double) * (size_t)k);
    if (evec != nullptr && k > 0) HIP_MultiVecGetToHost(ev, 0, k, evec, A->nrows);
    HIPGCG_Destroy(mv_ws, nullptr, nullptr, nev_max, bs, ops);
    ops->MultiVecDestroy(&ev, nev_max, ops);
    HIPGCG_Free(ops);
    OPS_HIP_Clear(ops);
    free(ops);
    HIPCCS_Destroy(&dA);
    if (dB) HIPCCS_Destroy(&dB);
    return 0;
}

}  // extern "C"

// GCG driver: a reentrant restatement of the reference eigensolver
// (src/ops_eig_sol_gcg.c:64-1562) with its inner solver BlockPCG
// (src/ops_lin_sol.c:140-465) and block modified Gram-Schmidt
// (src/ops_orth.c:45-118, 203-414).  It touches N-length data only
// through the operation table, so the same code runs over the HIP app and
// over the host dense table; every small matrix stays on the host as in
// the reference.  State lives in one heap object (ops->
// eigen_solver_workspace) instead of file-scope statics, so several
// tables can solve independently.
//
// Licence note: the phases, their step order, the parameter defaults and
// the monitor strings below follow the reference GCGE (xie-group/gcge,
// src/ops_eig_sol_gcg.c, src/ops_lin_sol.c, src/ops_orth.c; Yu Li et al.),
// which is distributed under the GNU Lesser General Public License 2.1.
// This file is a close restatement of that algorithm in new C++ (the FP64
// path must reproduce the reference's trajectory), so it is a work based
// on the reference and is distributed under the same LGPL-2.1 terms.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dense_eig.h"
#include "gcge_hip.h"

namespace gcge {
// Inner iterations the device FP32 BlockPCG of `ops` has run since it was
// installed (0 when the slot holds another solver); pcg32.hip.
int64_t pcg32_iterations(OPS *ops);
// r[:, sr:] = A x[:, sx:] - B x[:, sx:] diag(lam) over nc columns in one
// matrix-free pass when ops is the HIP table and A, B are 7-point stencil
// operators of one grid (false otherwise); app_hip.cpp.
bool hip_stencil_residual(OPS *ops, void *A, void *B, void **x, int sx, int nc,
                          const double *lam, void **r, int sr);
int hip_cholqr_pass(OPS *ops, void **x, int sx, int n, void *B, void **mv_ws, double zero_tol,
                    double rel_tol);
int hip_orth_coef_comb(OPS *ops, double *Z, int n, int sp, int np, double zero_tol,
                       void **V, int sv, void **ws, int dst);
}

namespace {

struct OrthParams {
    int block_size = -1;
    int max_reorth = 1;
    double zero_tol = 1e-14;
    double reorth_tol = 50 * DBL_EPSILON;  // src/ops_orth.c:404
};

struct Solver {
    // sizes and tolerances (EigenSolverSetup_GCG, src/ops_eig_sol_gcg.c:1565-1628)
    double gapMin = 0.01;
    int nevInit = 3, nevMax = 3, block_size = 1;
    double tol[2] = {1e-8, 1e-8};
    int numIterMax = 4;
    int print = 0, printtime = 0;
    int user_ls = 0;
    void **mv_ws[4] = {nullptr, nullptr, nullptr, nullptr};
    // algorithm parameters (static defaults of src/ops_eig_sol_gcg.c:1573-1603)
    int check_conv_max_num = 15;
    OrthParams orthX, orthP, orthW;
    // block orthonormalizations by CholQR2 and how many fell back to MGS
    int64_t cholqr_blocks = 0, cholqr_fallbacks = 0;
    int cg_max_iter = 40;
    double cg_rate = 1e-2, cg_tol = 1e-8;
    std::string cg_tol_type = "abs";
    int cg_auto_shift = 0;
    double cg_shift = 0.0;
    int compRR_min_num = -1;
    double compRR_min_gap = 0.01, compRR_tol = 1e-16;
    int eig_backend = -1;  // -1: automatic
    int orth_method = 0;   // GCGE_ORTH_MGS (the reference's), GCGE_ORTH_CHOLQR(_HOST)
    // run state
    OPS *ops = nullptr;
    void *A = nullptr, *B = nullptr;
    double sigma = 0.0;
    int sizeC = 0, sizeN = 0, sizeX = 0, sizeP = 0, sizeW = 0, sizeV = 0;
    int startN = 0, endN = 0, endX = 0, startP = 0, endP = 0, startW = 0, endW = 0;
    int indd = 0;
    std::vector<double> ss_eval, ss_matA, ss_evec;  // host small matrices
    std::vector<double> dbuf;                       // host scratch
    // statistics
    int numIter = 0;
    int64_t bpcg_iters = 0, linsolves = 0;
    int64_t pcg32_base = 0;  // the device BlockPCG's count when the solve began
    double times[10] = {0};
};

enum { T_INITX, T_CHECK, T_COMPP, T_RR, T_RRW, T_EIG, T_RV, T_COMPW, T_COMPX, T_LINSOL };

Solver *solver_of(OPS *ops)
{
    Solver *s = static_cast<Solver *>(ops->eigen_solver_workspace);
    if (s == nullptr) {
        fprintf(stderr, "gcge_hip: GCG used before HIPGCG_EigenSolverSetup\n");
        abort();
    }
    return s;
}

// ------------------------------------------------------------------ MGS
// Column-by-column B-orthonormalization of x[:, s:*e] (src/ops_orth.c:45-118).
void orth_self(void **x, int start_x, int *end_x, void *B, const OrthParams &p, void **mv_ws,
               OPS *ops)
{
    if (*end_x <= start_x) return;
    std::vector<double> r((size_t)(*end_x - start_x) + 1);
    int s[2], e[2];
    for (int k = start_x; k < *end_x; ++k) {
        s[0] = k; e[0] = *end_x;
        s[1] = k; e[1] = k + 1;
        ops->MultiVecQtAP('S', 'N', x, B, x, 0, s, e, r.data(), e[0] - s[0], mv_ws, ops);
        double rk = std::sqrt(r[0]);
        if (rk < p.zero_tol) {
            ops->Printf("r_[%d] = %6.4e\n", k, rk);
            if (k < *end_x - 1) {
                s[0] = *end_x - 1; e[0] = *end_x;
                s[1] = k; e[1] = k + 1;
                ops->MultiVecAxpby(1.0, x, 0.0, x, s, e, ops);
            }
            --k;
            --(*end_x);
            continue;
        }
        s[0] = k; e[0] = k + 1;
        s[1] = k; e[1] = k + 1;
        const double inv = 1.0 / rk;
        ops->MultiVecAxpby(0.0, nullptr, inv, x, s, e, ops);
        if (k < *end_x - 1) {
            const int len = *end_x - (k + 1);
            std::vector<double> coef((size_t)len);
            for (int i = 0; i < len; ++i) coef[i] = -inv * r[1 + i];
            double one = 1.0;
            s[0] = k; e[0] = k + 1;
            s[1] = k + 1; e[1] = *end_x;
            ops->MultiVecLinearComb(x, x, 0, s, e, coef.data(), 1, &one, 0, ops);
            for (int it = 1; it < p.max_reorth - 1; ++it) {
                s[0] = k + 1; e[0] = *end_x;
                s[1] = k; e[1] = k + 1;
                ops->MultiVecQtAP('S', 'N', x, B, x, 0, s, e, coef.data(), len, mv_ws, ops);
                double amax = 0.0;
                for (int i = 0; i < len; ++i) {
                    coef[i] = -coef[i];
                    amax = std::max(amax, std::fabs(coef[i]));
                }
                s[0] = k; e[0] = k + 1;
                s[1] = k + 1; e[1] = *end_x;
                ops->MultiVecLinearComb(x, x, 0, s, e, coef.data(), 1, &one, 0, ops);
                if (amax < p.reorth_tol) break;
            }
        }
    }
}

// Removes from x[:, start_x:end_x] its components along x[:, :start_x] in
// the B inner product: block classical Gram-Schmidt passes until the
// largest coefficient is below reorth_tol, at most 1 + max_reorth
// (src/ops_orth.c:231-269), or at most max_passes when that is positive.
void remove_previous(void **x, int start_x, int end_x, void *B, const OrthParams &p,
                     void **mv_ws, OPS *ops, int max_passes = 0)
{
    if (start_x <= 0 || end_x <= start_x) return;
    int s[2], e[2];
    double one = 1.0;
    s[0] = 0; e[0] = start_x;
    s[1] = start_x; e[1] = end_x;
    const int64_t len = (int64_t)(e[1] - s[1]) * (e[0] - s[0]);
    std::vector<double> coef((size_t)len);
    int passes = 1 + p.max_reorth;
    if (max_passes > 0) passes = std::min(passes, max_passes);
    for (int it = 0; it < passes; ++it) {
        ops->MultiVecQtAP('S', 'N', x, B, x, 0, s, e, coef.data(), e[0] - s[0], mv_ws, ops);
        double amax = 0.0;
        for (int64_t i = 0; i < len; ++i) {
            coef[i] = -coef[i];
            amax = std::max(amax, std::fabs(coef[i]));
        }
        ops->MultiVecLinearComb(x, x, 0, s, e, coef.data(), e[0] - s[0], &one, 0, ops);
        if (amax < p.reorth_tol) break;
    }
}

void mgs_blocks(void **x, int start_x, int *end_x, void *B, const OrthParams &p, void **mv_ws,
                OPS *ops);

// Cholesky G = R^T R of an n x n symmetric positive definite matrix
// (upper R, column-major).  Returns false when a pivot is not above
// rel_tol times the largest diagonal entry (numerically dependent columns).
bool cholesky_upper(int n, const std::vector<double> &G, std::vector<double> &R, double rel_tol)
{
    R.assign((size_t)n * n, 0.0);
    double dmax = 0.0;
    for (int i = 0; i < n; ++i) dmax = std::max(dmax, G[(size_t)n * i + i]);
    if (!(dmax > 0.0)) return false;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i <= j; ++i) {
            double s = G[(size_t)n * j + i];
            for (int k = 0; k < i; ++k) s -= R[(size_t)n * i + k] * R[(size_t)n * j + k];
            if (i == j) {
                if (!(s > rel_tol * dmax)) return false;
                R[(size_t)n * j + j] = std::sqrt(s);
            } else {
                R[(size_t)n * j + i] = s / R[(size_t)n * i + i];
            }
        }
    }
    return true;
}

// Block orthonormalization of x[:, start_x:*end_x] in the B inner product:
// the previous columns are removed as in MGS, then Cholesky QR is applied
// twice to the whole block (CholQR2): G = X^T B X, scaled to unit diagonal
// G' = D^{-1} G D^{-1} (D = diag(G)^{1/2}, so columns of very different
// norms -- the corrections of nearly converged pairs are tiny -- do not
// look dependent), G' = R^T R, X <- X D^{-1} R^{-1}.  The same subspace as
// the reference's column-by-column MGS with a few block products instead
// of one B-product and two passes per column.  A column whose B-norm is
// below zero_tol (which MGS drops) or a numerically dependent block sends
// the block to the column-by-column MGS.
//
// The first pass scales the block's remainders along the earlier columns
// (of order eps after the removal) up: by a column's cancellation when it
// lost digits to the earlier columns, and by cond(R) when the block's own
// columns are nearly dependent.  The second pass orthonormalizes only
// within the block, so the earlier columns are removed once more between
// the passes, as MGS reorthogonalizes.  Without it, blocks of the mixed
// path meeting the degenerate clusters of 3D Laplacians lost
// B-orthogonality: a 16^3 solve stalled on spurious near-zero Ritz values
// and a 48^3 solve returned vectors with max|X^T B X - I| = 2e-8.  It costs
// one Gram and one LinearComb per block (1.5 % of the 128^3 mixed solve).
//
// With device set and ops the HIP table, each pass's Gram, factorization,
// inverse and product stay on the device (hip_cholqr_pass): the same
// tests and the same X D^{-1} R^{-1}, with the factorization's own
// rounding (right-looking, R^{-1} from the inverse of R^T by forward
// substitution).
void cholqr(void **x, int start_x, int *end_x, void *B, const OrthParams &p, void **mv_ws,
            OPS *ops, int64_t *fallbacks, bool device)
{
    if (*end_x <= start_x) return;
    // two block passes ("twice is enough" for classical Gram-Schmidt: the
    // second pass leaves components of order eps * ||x||, Giraud, Langou
    // and Rozloznik 2005); the reference's absolute test on the largest
    // coefficient (50 eps) would run a third pass on unit-size columns
    remove_previous(x, start_x, *end_x, B, p, mv_ws, ops, 2);
    const int n = *end_x - start_x;
    std::vector<double> G((size_t)n * n), R, Rinv((size_t)n * n), dinv((size_t)n);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) remove_previous(x, start_x, *end_x, B, p, mv_ws, ops, 1);
        const int done =
            device ? gcge::hip_cholqr_pass(ops, x, start_x, n, B, mv_ws, p.zero_tol, 1e-10) : -1;
        if (done == 0) {
            *fallbacks += 1;
            mgs_blocks(x, start_x, end_x, B, p, mv_ws, ops);
            return;
        }
        if (done == 1) continue;
        int s[2] = {start_x, start_x}, e[2] = {*end_x, *end_x};
        ops->MultiVecQtAP('S', 'S', x, B, x, 0, s, e, G.data(), n, mv_ws, ops);
        bool ok = true;
        for (int i = 0; i < n && ok; ++i) {
            const double d = std::sqrt(std::max(G[(size_t)n * i + i], 0.0));
            ok = d >= p.zero_tol;
            dinv[i] = ok ? 1.0 / d : 0.0;
        }
        if (ok) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) G[(size_t)n * j + i] *= dinv[i] * dinv[j];
            ok = cholesky_upper(n, G, R, 1e-10);
        }
        if (!ok) {
            *fallbacks += 1;
            mgs_blocks(x, start_x, end_x, B, p, mv_ws, ops);
            return;
        }
        // R^{-1}, upper triangular, by back substitution column by column
        std::fill(Rinv.begin(), Rinv.end(), 0.0);
        for (int j = 0; j < n; ++j) {
            Rinv[(size_t)n * j + j] = 1.0 / R[(size_t)n * j + j];
            for (int i = j - 1; i >= 0; --i) {
                double t = 0.0;
                for (int k = i + 1; k <= j; ++k) t += R[(size_t)n * k + i] * Rinv[(size_t)n * j + k];
                Rinv[(size_t)n * j + i] = -t / R[(size_t)n * i + i];
            }
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i <= j; ++i) Rinv[(size_t)n * j + i] *= dinv[i];
        // X <- X D^{-1} R^{-1} through the workspace (the product cannot run in place)
        int s1[2] = {start_x, 0}, e1[2] = {*end_x, n};
        ops->MultiVecLinearComb(x, mv_ws, 0, s1, e1, Rinv.data(), n, nullptr, 0, ops);
        int s2[2] = {0, start_x}, e2[2] = {n, *end_x};
        ops->MultiVecAxpby(1.0, mv_ws, 0.0, x, s2, e2, ops);
    }
}

// Block MGS of x[:, start_x:*end_x] against x[:, :start_x] and itself
// (src/ops_orth.c:203-394).  mv_ws receives B times the current block.
void mgs(void **x, int start_x, int *end_x, void *B, const OrthParams &p, void **mv_ws,
         OPS *ops)
{
    if (*end_x <= start_x) return;
    remove_previous(x, start_x, *end_x, B, p, mv_ws, ops);
    mgs_blocks(x, start_x, end_x, B, p, mv_ws, ops);
}

// The self-orthonormalization part of MGS: blocks of half the range, each
// orthonormalized column by column and then removed from the rest
// (src/ops_orth.c:271-368).
void mgs_blocks(void **x, int start_x, int *end_x, void *B, const OrthParams &p, void **mv_ws,
                OPS *ops)
{
    int s[2], e[2];
    double one = 1.0;
    std::vector<double> coef;
    if (*end_x <= start_x) return;
    int init_start = start_x, init_end;
    int bsz = p.block_size;
    if (bsz <= 0) bsz = std::max((*end_x - init_start) / 2, 2);
    bsz = std::min(bsz, *end_x - init_start);
    while (bsz > 0) {
        s[1] = init_start;
        e[1] = s[1] + bsz;
        orth_self(x, s[1], &e[1], B, p, mv_ws, ops);
        init_end = e[1];
        int length = bsz - (e[1] - s[1]);
        length = std::min(length, *end_x - e[1] - length);
        if (length > 0) {
            e[0] = *end_x; s[0] = e[0] - length;
            s[1] = init_end; e[1] = s[1] + length;
            ops->MultiVecAxpby(1.0, x, 0.0, x, s, e, ops);
        }
        *end_x = *end_x - (bsz - (init_end - init_start));
        if (init_end < *end_x && init_start < init_end) {
            const int nb = init_end - init_start, nq = *end_x - init_end;
            coef.resize((size_t)nb * nq);
            for (int it = 0; it < 1 + p.max_reorth; ++it) {
                if (B != nullptr && it > 0) {
                    s[0] = init_end; e[0] = *end_x;
                    s[1] = 0; e[1] = nb;
                    ops->MultiVecQtAP('S', 'T', x, nullptr, mv_ws, 0, s, e, coef.data(), nb,
                                      mv_ws, ops);
                } else {
                    s[0] = init_end; e[0] = *end_x;
                    s[1] = init_start; e[1] = init_end;
                    ops->MultiVecQtAP('S', 'T', x, B, x, 0, s, e, coef.data(), nb, mv_ws, ops);
                }
                double amax = 0.0;
                for (size_t i = 0; i < coef.size(); ++i) {
                    coef[i] = -coef[i];
                    amax = std::max(amax, std::fabs(coef[i]));
                }
                s[0] = init_start; e[0] = init_end;
                s[1] = init_end; e[1] = *end_x;
                ops->MultiVecLinearComb(x, x, 0, s, e, coef.data(), nb, &one, 0, ops);
                if (amax < p.reorth_tol) break;
            }
        }
        init_start = init_end;
        bsz = std::min(bsz, *end_x - init_start);
    }
}

// ------------------------------------------------------------- BlockPCG
// y = (A + sigma B) x using z[:, s:] as scratch (src/ops_eig_sol_gcg.c:64-97)
void shifted_apply(Solver &S, void **x, void **y, int *start, int *end, void **z, int zs)
{
    OPS *ops = S.ops;
    ops->MatDotMultiVec(S.A, x, y, start, end, ops);
    if (S.sigma != 0.0) {
        if (S.B == nullptr) {
            ops->MultiVecAxpby(S.sigma, x, 1.0, y, start, end, ops);
        } else {
            int st[2], et[2];
            st[0] = start[0]; et[0] = end[0];
            st[1] = zs; et[1] = zs + end[0] - start[0];
            ops->MatDotMultiVec(S.B, x, z, st, et, ops);
            st[0] = zs; et[0] = zs + end[0] - start[0];
            st[1] = start[1]; et[1] = end[1];
            ops->MultiVecAxpby(S.sigma, z, 1.0, y, st, et, ops);
        }
    }
}

// Per-column CG (src/ops_lin_sol.c:140-437); workspaces r, p, w are
// S.mv_ws[1..3] as set up by ComputeW (src/ops_eig_sol_gcg.c:611-616).
void block_pcg(Solver &S, void **mv_b, void **mv_x, int *start_bx, int *end_bx)
{
    OPS *ops = S.ops;
    void **r = S.mv_ws[1], **p = S.mv_ws[2], **w = S.mv_ws[3];
    const int n0 = end_bx[0] - start_bx[0];
    const double rate = S.cg_rate, tol = S.cg_tol;
    std::vector<double> norm_b((size_t)n0, 1.0), rho1((size_t)n0), rho2((size_t)n0),
        pTw((size_t)n0), init_res((size_t)n0), last_res((size_t)n0);
    std::vector<int> unconv((size_t)n0), block((size_t)n0 + 1);
    int s[2], e[2];
    if (S.cg_tol_type == "rel") {
        s[0] = start_bx[0]; e[0] = end_bx[0];
        s[1] = start_bx[0]; e[1] = end_bx[0];
        ops->MultiVecInnerProd('D', mv_b, mv_b, 0, s, e, norm_b.data(), 1, ops);
        for (auto &v : norm_b) v = std::sqrt(v);
    }
    s[0] = start_bx[1]; e[0] = end_bx[1];
    s[1] = 0; e[1] = n0;
    shifted_apply(S, mv_x, r, s, e, p, 0);
    s[0] = start_bx[0]; e[0] = end_bx[0];
    s[1] = 0; e[1] = n0;
    ops->MultiVecAxpby(1.0, mv_b, -1.0, r, s, e, ops);
    s[0] = 0; e[0] = n0;
    s[1] = 0; e[1] = n0;
    ops->MultiVecInnerProd('D', r, r, 0, s, e, rho2.data(), 1, ops);
    for (int i = 0; i < n0; ++i) init_res[i] = std::sqrt(rho2[i]);
    int nu = 0;
    for (int i = 0; i < n0; ++i) {
        if (init_res[i] > tol * norm_b[i]) {
            unconv[nu] = i;
            rho2[nu] = rho2[i];
            ++nu;
        }
    }
    int niter = 0;
    while (niter < S.cg_max_iter && nu > 0) {
        int nblock = 0;
        block[nblock++] = 0;
        for (int i = 1; i < nu; ++i)
            if (unconv[i] - unconv[i - 1] > 1) block[nblock++] = i;
        block[nblock] = nu;
        double *dst = pTw.data();
        for (int b = 0; b < nblock; ++b) {
            const int len = block[b + 1] - block[b];
            for (int c = block[b]; c < block[b + 1]; ++c) {
                const double beta = (niter == 0) ? 0.0 : rho2[c] / rho1[c];
                s[0] = unconv[c]; e[0] = unconv[c] + 1;
                s[1] = unconv[c]; e[1] = unconv[c] + 1;
                ops->MultiVecAxpby(1.0, r, beta, p, s, e, ops);
            }
            s[0] = unconv[block[b]]; e[0] = unconv[block[b + 1] - 1] + 1;
            s[1] = s[0]; e[1] = e[0];
            shifted_apply(S, p, w, s, e, mv_b, start_bx[0]);
            ops->MultiVecLocalInnerProd('D', p, w, 0, s, e, dst, 1, ops);
            dst += len;
        }
        for (int i = 0; i < nu; ++i) rho1[i] = rho2[i];
        dst = rho2.data();
        for (int b = 0; b < nblock; ++b) {
            const int len = block[b + 1] - block[b];
            for (int c = block[b]; c < block[b + 1]; ++c) {
                const double alpha = rho2[c] / pTw[c];
                s[0] = unconv[c]; e[0] = unconv[c] + 1;
                s[1] = start_bx[1] + unconv[c]; e[1] = s[1] + 1;
                ops->MultiVecAxpby(alpha, p, 1.0, mv_x, s, e, ops);
                s[0] = unconv[c]; e[0] = unconv[c] + 1;
                s[1] = unconv[c]; e[1] = unconv[c] + 1;
                ops->MultiVecAxpby(-alpha, w, 1.0, r, s, e, ops);
            }
            s[0] = unconv[block[b]]; e[0] = unconv[block[b + 1] - 1] + 1;
            s[1] = s[0]; e[1] = e[0];
            ops->MultiVecLocalInnerProd('D', r, r, 0, s, e, dst, 1, ops);
            dst += len;
        }
        for (int i = 0; i < nu; ++i) last_res[unconv[i]] = std::sqrt(rho2[i]);
        const int pre = nu;
        nu = 0;
        for (int i = 0; i < pre; ++i) {
            const int c = unconv[i];
            if (last_res[c] > rate * init_res[c] && last_res[c] > tol * norm_b[c]) {
                unconv[nu] = c;
                rho1[nu] = rho1[i];
                rho2[nu] = rho2[i];
                ++nu;
            }
        }
        ++niter;
    }
    S.bpcg_iters += niter;
    S.linsolves += 1;
}

void block_pcg_entry(void *, void **b, void **x, int *start, int *end, OPS *ops)
{
    block_pcg(*solver_of(ops), b, x, start, end);
}

// ------------------------------------------------------------------ phases
struct Timer {
    Solver &S;
    int slot;
    double t0;
    Timer(Solver &s, int k) : S(s), slot(k), t0(s.ops->GetWtime()) {}
    ~Timer() { S.times[slot] += S.ops->GetWtime() - t0; }
};

int nV(const Solver &S) { return S.sizeV - S.sizeC; }

// B-orthonormalization of the new W block or the initial X block with the
// selected method (the reference always uses MGS)
void orthonormalize(Solver &S, void **V, int start, int *end, const OrthParams &p, void **ws)
{
    if (S.orth_method == GCGE_ORTH_CHOLQR || S.orth_method == GCGE_ORTH_CHOLQR_HOST)
    {
        S.cholqr_blocks += 1;
        cholqr(V, start, end, S.B, p, ws, S.ops, &S.cholqr_fallbacks,
               S.orth_method == GCGE_ORTH_CHOLQR);
    }
    else
        mgs(V, start, end, S.B, p, ws, S.ops);
}

void initialize_x(Solver &S, void **V, void **ritz, int nevGiven)
{
    Timer t(S, T_INITX);
    OPS *ops = S.ops;
    int s[2] = {0, 0}, e[2] = {nevGiven, nevGiven};
    ops->MultiVecAxpby(1.0, ritz, 0.0, V, s, e, ops);
    mgs(V, 0, &nevGiven, S.B, S.orthX, ritz, ops);
    ops->MultiVecSetRandomValue(V, nevGiven, S.sizeX, ops);
    S.endX = S.sizeX;
    orthonormalize(S, V, nevGiven, &S.endX, S.orthX, ritz);
    if (S.endX != S.sizeX) {
        fprintf(stderr, "gcge_hip: initial block lost rank (%d of %d)\n", S.endX, S.sizeX);
        abort();
    }
}

void ritz_vec(Solver &S, void **ritz, void **V)
{
    Timer t(S, T_RV);
    int s[2] = {S.startN, S.startN}, e[2] = {S.endW, S.endX};
    S.ops->MultiVecLinearComb(V, ritz, 0, s, e, S.ss_evec.data(), nV(S), nullptr, 0, S.ops);
}

int check_convergence(Solver &S, void **ritz, int num_check, std::vector<int> &offset)
{
    Timer t(S, T_CHECK);
    OPS *ops = S.ops;
    void **ws0 = S.mv_ws[1], **ws1 = S.mv_ws[2];
    int s[2] = {S.startN, 0}, e[2] = {S.startN + num_check, num_check};
    if (!gcge::hip_stencil_residual(ops, S.A, S.B, ritz, S.startN, num_check,
                                    S.ss_eval.data() + S.startN, ws0, 0)) {
        ops->MatDotMultiVec(S.A, ritz, ws0, s, e, ops);
        ops->MatDotMultiVec(S.B, ritz, ws1, s, e, ops);
        ops->MultiVecLinearComb(nullptr, ws1, 0, s, e, nullptr, 0, S.ss_eval.data() + S.startN,
                                1, ops);
        s[0] = 0; e[0] = num_check;
        s[1] = 0; e[1] = num_check;
        ops->MultiVecAxpby(-1.0, ws1, 1.0, ws0, s, e, ops);
    }
    s[0] = 0; e[0] = num_check;
    s[1] = 0; e[1] = num_check;
    std::vector<double> res((size_t)num_check);
    ops->MultiVecInnerProd('D', ws0, ws0, 0, s, e, res.data(), 1, ops);
    for (auto &v : res) v = std::sqrt(v);
    const double *ev = S.ss_eval.data();
    auto unconv = [&](int i) {
        return res[i] > S.tol[0] || res[i] > std::fabs(ev[S.startN + i]) * S.tol[1];
    };
    int idx = 0;
    for (; idx < num_check; ++idx) {
        if (unconv(idx)) {
            if (S.print)
                ops->Printf("%3d EPS nconv=%d first unconverged value (error(abs,rel)) %6.5e "
                            "(%6.5e, %6.5e)\n",
                            S.indd + 1, S.startN, ev[S.startN + idx], res[idx],
                            res[idx] / std::fabs(ev[S.startN + idx]));
            break;
        }
    }
    for (; idx > 0; --idx) {
        const double a = ev[S.startN + idx - 1];
        if (std::fabs((a - ev[S.startN + idx]) / a) > S.gapMin) break;
    }
    const int nev_conv = S.sizeC + idx;
    // unconverged ranges (src/ops_eig_sol_gcg.c:268-307)
    offset.assign(2 * (size_t)S.sizeN + 3, 0);
    int state = 1, nu = 0;
    for (int i = 0; i < num_check; ++i) {
        if (unconv(i)) {
            if (state) {
                offset[offset[0] * 2 + 1] = S.startN + i;
                state = 0;
            }
            ++nu;
            if (nu == S.sizeN) {
                offset[offset[0] * 2 + 2] = S.startN + i + 1;
                ++offset[0];
                break;
            }
        } else if (!state) {
            offset[offset[0] * 2 + 2] = S.startN + i;
            ++offset[0];
            state = 1;
        }
    }
    if (nu < S.sizeN) {
        if (state == 1) offset[offset[0] * 2 + 1] = S.startN + num_check;
        int eo = S.startN + num_check + S.sizeN - nu;
        offset[offset[0] * 2 + 2] = std::min(eo, S.endX);
        if (!(offset[offset[0] * 2 + 1] < offset[offset[0] * 2 + 2])) {
            fprintf(stderr, "gcge_hip: empty unconverged range\n");
            abort();
        }
        ++offset[0];
    }
    return nev_conv;
}

void compute_p(Solver &S, void **V, const std::vector<int> &offset)
{
    Timer t(S, T_COMPP);
    OPS *ops = S.ops;
    const int n = nV(S);
    double *Z = S.ss_evec.data();
    int bsz = 0;
    for (int i = 0; i < offset[0]; ++i) {
        const int o1 = offset[2 * i + 1], o2 = offset[2 * i + 2];
        memmove(Z + (size_t)n * (S.sizeX - S.sizeC + bsz), Z + (size_t)n * (o1 - S.sizeC),
                sizeof(double) * n * (o2 - o1));
        bsz += o2 - o1;
    }
    S.sizeP = bsz;
    for (int i = 0; i < offset[0]; ++i) {
        const int o1 = offset[2 * i + 1], o2 = offset[2 * i + 2];
        for (int c = 0; c < S.sizeP; ++c)
            memset(Z + (size_t)n * (S.sizeX - S.sizeC + c) + (o1 - S.sizeC), 0,
                   sizeof(double) * (o2 - o1));
    }
    int startP = S.sizeX - S.sizeC, endP = startP + S.sizeP;
    // With the block orthonormalization on the device, coefficient blocks of
    // 4 M multiply-adds and more (order 460 by 120 columns: 19 M) are
    // orthonormalized there too, by Cholesky QR instead of the MGS below
    // (same span; an MGS on the host costs ~5 ms at that size, the device
    // steps ~0.6 ms); dependent columns fall back to the MGS.
    if (S.orth_method == GCGE_ORTH_CHOLQR && S.sizeP > 0 && S.endW - S.startN == n &&
        (double)n * S.sizeP * endP >= 4e6 &&
        gcge::hip_orth_coef_comb(ops, Z, n, startP, S.sizeP, S.orthP.zero_tol, V, S.startN,
                                 S.mv_ws[1], startP + S.sizeC) == 1) {
        S.startP = startP + S.sizeC;
        S.endP = endP + S.sizeC;
        S.sizeP = S.endP - S.startP;
        return;
    }
    // small MGS of the P coefficients in the identity inner product
    gcge_dense_mv vecP{Z, n, endP, n};
    std::vector<double> wsbuf((size_t)n * (S.sizeP > 0 ? S.sizeP : 1));
    gcge_dense_mv vecW{wsbuf.data(), n, S.sizeP, n};
    // host table for the coefficients: ops->lapack_ops, or the table itself
    // when it is the host dense table (whose lapack_ops is NULL, as
    // OPS_LAPACK_Set leaves it)
    OPS *host_ops = ops->lapack_ops ? ops->lapack_ops : ops;
    mgs(reinterpret_cast<void **>(&vecP), startP, &endP, nullptr, S.orthP,
        reinterpret_cast<void **>(&vecW), host_ops);
    S.startP = startP + S.sizeC;
    S.endP = endP + S.sizeC;
    S.sizeP = S.endP - S.startP;
    void **ws0 = S.mv_ws[1];
    int s[2] = {S.startN, 0}, e[2] = {S.endW, S.sizeP};
    ops->MultiVecLinearComb(V, ws0, 0, s, e, Z + (size_t)n * (S.sizeX - S.sizeC), n, nullptr, 0,
                            ops);
    s[0] = 0; e[0] = S.sizeP;
    s[1] = S.startP; e[1] = S.endP;
    ops->MultiVecAxpby(1.0, ws0, 0.0, V, s, e, ops);
}

void compute_w(Solver &S, void **V, double *ss_eval, void **ritz, const std::vector<int> &offset)
{
    Timer t(S, T_COMPW);
    OPS *ops = S.ops;
    void **b = ritz;
    double sigma = 0.0;
    if (S.cg_auto_shift == 1)
        sigma = -ss_eval[S.sizeC] + (ss_eval[S.sizeC + 1] - ss_eval[S.sizeC]) * 0.01;
    S.sigma = S.cg_shift + sigma;
    sigma = S.sigma;
    if (!(S.cg_auto_shift == 0 || S.user_ls == 0)) {
        fprintf(stderr, "gcge_hip: a user linear solver cannot be combined with autoshift\n");
        abort();
    }
    int bsz = 0;
    S.startW = S.endP;
    int s[2], e[2];
    for (int i = 0; i < offset[0]; ++i) {
        const int o1 = offset[2 * i + 1], o2 = offset[2 * i + 2], len = o2 - o1;
        s[0] = o1; e[0] = o2;
        s[1] = S.startW + bsz; e[1] = s[1] + len;
        ops->MultiVecAxpby(1.0, ritz, 0.0, V, s, e, ops);
        s[0] = o1; e[0] = o2;
        s[1] = offset[1] + bsz; e[1] = s[1] + len;
        ops->MatDotMultiVec(S.B, V, b, s, e, ops);
        std::vector<double> scal((size_t)len);
        for (int k = 0; k < len; ++k) scal[k] = ss_eval[o1 + k] + sigma;
        ops->MultiVecLinearComb(nullptr, b, 0, s, e, nullptr, 0, scal.data(), 1, ops);
        bsz += len;
    }
    S.endW = S.startW + bsz;
    s[0] = offset[1]; e[0] = s[0] + bsz;
    s[1] = S.startW; e[1] = S.endW;
    {
        Timer tl(S, T_LINSOL);
        if (S.user_ls == 0 || S.user_ls == 2) {
            block_pcg(S, b, V, s, e);
        } else {
            ops->MultiLinearSolver(S.A, b, V, s, e, ops);
            S.linsolves += 1;
        }
    }
    orthonormalize(S, V, S.startW, &S.endW, S.orthW, S.mv_ws[1]);
    S.sizeW = S.endW - S.startW;
}

void dense_eigh(Solver &S, int n, double *M, double *w, double *Z)
{
    Timer t(S, T_EIG);
    int backend = S.eig_backend;
    // the host solver wins below ~128 (no launches; measured 2.5 ms at 90,
    // 10 ms at 150, against 5-10 ms for the device solver at 220-460)
    if (backend < 0) backend = (n >= 128) ? GCGE_EIG_DEVICE : GCGE_EIG_HOST;
    int info;
    if (backend == GCGE_EIG_DEVICE)
        info = gcge::eigh_device(n, M, n, w, Z, n);
    else
        info = gcge::eigh_host(n, M, n, w, Z, n);
    if (info != 0) {
        fprintf(stderr, "gcge_hip: dense eigensolver failed (info %d, n %d)\n", info, n);
        abort();
    }
}

// The returned values of the k converged pairs as their vectors' Rayleigh
// quotients x^T A x / x^T B x, and the vectors scaled to unit B-norm (the
// block-orthonormalization path).  Over many iterations the converged
// (locked) vectors can drift from B-orthonormality, and a Ritz value of the
// projected problem then differs from its vector's Rayleigh quotient by
// about that drift (a served 48^3 solve reached max|X^T B X - I| = 2e-8
// with Ritz values 2e-9 from their Rayleigh quotients, which were exact to
// 2e-14).  The vectors keep their directions, so each residual
// ||A x - theta B x|| only drops (theta is the quotient that minimizes it
// in the B^{-1} norm).  V (no longer needed) is the workspace.
void final_rayleigh_quotients(Solver &S, void **X, int k, double *eval, void **V)
{
    Timer t(S, T_RV);
    OPS *ops = S.ops;
    if (k <= 0) return;
    std::vector<double> xa((size_t)k), xb((size_t)k), sc((size_t)k);
    int s[2] = {0, 0}, e[2] = {k, k};
    ops->MatDotMultiVec(S.B, X, V, s, e, ops);
    ops->MultiVecInnerProd('D', X, V, 0, s, e, xb.data(), 1, ops);
    ops->MatDotMultiVec(S.A, X, V, s, e, ops);
    ops->MultiVecInnerProd('D', X, V, 0, s, e, xa.data(), 1, ops);
    for (int j = 0; j < k; ++j) {
        if (!(xb[j] > 0.0)) return;  // not a B-positive vector: keep the iteration's
        sc[j] = 1.0 / std::sqrt(xb[j]);
    }
    for (int j = 0; j < k; ++j) eval[j] = xa[j] / xb[j];
    ops->MultiVecLinearComb(nullptr, X, 0, s, e, nullptr, 0, sc.data(), 1, ops);
}

void rayleigh_ritz(Solver &S, int nev_conv, void **V)
{
    Timer t(S, T_RR);
    OPS *ops = S.ops;
    std::vector<double> PtAP;
    const int sX = S.sizeX, sP = S.sizeP, sW = S.sizeW;
    if (sP > 0) {
        // P^T (old projected matrix) P from the coefficients, lower triangle read
        const int n = nV(S);
        PtAP.assign((size_t)sP * sP, 0.0);
        std::vector<double> ws((size_t)n * sP);
        double *Pc = S.ss_evec.data() + (size_t)n * (sX - S.sizeC);
        ops->DenseMatQtAP('L', 'S', n, n, sP, sP, 1.0, Pc, n, S.ss_matA.data(), n, Pc, n, 0.0,
                          PtAP.data(), sP, ws.data());
    }
    S.sizeV = sX + sP + sW;
    S.startN += nev_conv - S.sizeC;
    S.endN += nev_conv - S.sizeC;
    S.endN = std::min(S.endN, S.endX);
    S.sizeN = S.endN - S.startN;
    S.sizeC = nev_conv;
    const int n = nV(S), sC = S.sizeC;
    std::vector<double> M((size_t)n * n, 0.0);
    if (sW > 0) {
        Timer tw(S, T_RRW);
        int s[2] = {S.startN, S.startW}, e[2] = {S.endW, S.endW};
        const int off = sX + sP - sC;
        ops->MultiVecQtAP('S', 'N', V, S.A, V, 0, s, e, M.data() + (size_t)n * off, n,
                          S.mv_ws[1], ops);
        for (int j = 0; j < sW; ++j)
            for (int i = 0; i < off; ++i) M[(size_t)n * i + off + j] = M[(size_t)n * (off + j) + i];
    }
    if (sX == S.sizeV) {
        int length = sX - sC;
        int b = std::min(S.block_size, length);
        int s[2] = {sC, sC}, e[2] = {sX, sC + b};
        double *dst = M.data();
        while (length) {
            ops->MultiVecQtAP('S', 'N', V, S.A, V, 0, s, e, dst, n, S.mv_ws[1], ops);
            dst += (size_t)n * b;
            length -= b;
            b = std::min(b, length);
            s[1] = e[1];
            e[1] = s[1] + b;
        }
    } else {
        for (int k = 0; k < sX - sC; ++k) M[(size_t)n * k + k] = S.ss_eval[sC + k];
        for (int j = 0; j < sP; ++j)
            for (int i = 0; i < sP; ++i)
                M[(size_t)n * (sX - sC + j) + (sX - sC + i)] = PtAP[(size_t)sP * j + i];
    }
    // the reference's dsyevx reads the upper triangle (UPLO = 'U'); keep a
    // symmetric copy built from it for the next step's P^T A P
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) M[(size_t)n * j + i] = M[(size_t)n * i + j];
    if (S.cg_shift != 0.0)
        for (int k = 0; k < n; ++k) M[(size_t)n * k + k] += S.cg_shift;
    S.ss_matA = M;
    // n x n eigenvectors plus block_size spare columns: the P step copies
    // up to sizeN coefficient columns after column sizeX - sizeC, which can
    // pass column n when orthogonalization dropped W columns (the reference
    // spills into the workspace that follows ss_evec)
    S.ss_evec.assign((size_t)n * (n + S.block_size), 0.0);
    std::vector<double> w((size_t)n);
    dense_eigh(S, n, M.data(), w.data(), S.ss_evec.data());
    for (int k = 0; k < n; ++k) S.ss_eval[sC + k] = w[k] - S.cg_shift;
    if (S.cg_shift != 0.0)
        for (int k = 0; k < n; ++k) S.ss_matA[(size_t)n * k + k] -= S.cg_shift;
}

void gcg(void *A, void *B, double *eval, void **evec, int nevGiven, int *nevConv, OPS *ops)
{
    Solver &S = *solver_of(ops);
    S.ops = ops;
    S.A = A;
    S.B = B;
    for (double &t : S.times) t = 0.0;
    S.bpcg_iters = 0;
    S.pcg32_base = gcge::pcg32_iterations(ops);
    S.cholqr_blocks = 0;
    S.cholqr_fallbacks = 0;
    S.linsolves = 0;
    const int bs = S.block_size, nevInit = S.nevInit, nevMax = S.nevMax;
    if (!(nevInit >= nevGiven && nevInit <= nevMax && (nevInit >= 3 * bs || nevInit == nevMax) &&
          nevMax >= *nevConv + bs && nevMax <= *nevConv + nevInit)) {
        fprintf(stderr,
                "gcge_hip: sizes violate the driver's workspace rules "
                "(nevInit %d nevMax %d block %d nev %d)\n",
                nevInit, nevMax, bs, *nevConv);
        abort();
    }
    int numIterMax = S.numIterMax;
    S.sizeC = 0;
    S.sizeN = bs;
    S.sizeX = nevInit;
    S.sizeP = 0;
    S.sizeW = 0;
    S.sizeV = S.sizeX;
    S.startN = 0;
    S.endN = bs;
    S.endX = S.sizeX;
    S.startP = S.endX;
    S.endP = S.startP;
    S.startW = S.endP;
    S.endW = S.startW;
    S.indd = 0;
    void **V = S.mv_ws[0], **ritz = evec;
    S.ss_eval.assign((size_t)nevMax + 2 * bs, 1.0);
    std::vector<int> offW, offP;

    initialize_x(S, V, ritz, nevGiven);
    rayleigh_ritz(S, 0, V);
    for (int i = S.sizeV; i < nevMax + 2 * bs; ++i) S.ss_eval[i] = S.ss_eval[S.sizeV - 1];
    ritz_vec(S, ritz, V);

    *nevConv = std::min(*nevConv, nevMax);
    const int nev0 = *nevConv;
    *nevConv = 0;
    int nev = nevInit < nevMax ? 2 * bs : nev0;
    nev = std::min(nev, nev0);
    int numIter = 0;
    do {
        int numCheck;
        if (numIter <= 0)
            numCheck = 0;
        else
            numCheck = (S.startN + S.sizeN < S.endX) ? S.sizeN : S.endX - S.startN;
        numCheck = std::min(numCheck, S.check_conv_max_num);
        if (numCheck == 0) numCheck = 1;
        *nevConv = check_convergence(S, ritz, numCheck, offW);
        if (*nevConv >= nev) {
            if (*nevConv >= nev0) break;
            nev += S.sizeP + S.sizeW;
            nev = std::min(nev, nev0);
            S.sizeX += S.sizeP + S.sizeW;
            S.sizeX = std::min(S.sizeX, nevMax);
            int s[2] = {S.startN, S.endX}, e[2] = {S.endW, S.sizeX};
            ops->MultiVecLinearComb(V, ritz, 0, s, e,
                                    S.ss_evec.data() + (size_t)nV(S) * (S.endX - S.sizeC), nV(S),
                                    nullptr, 0, ops);
            S.sizeP = 0;
            S.sizeW = 0;
            S.sizeV = S.sizeX;
            S.startP = S.endX;
            S.endP = S.startP;
            S.startW = S.endP;
            S.endW = S.startW;
            S.endX = S.sizeX;
            S.endN = std::min(S.startN + bs, S.endX);
            S.sizeN = S.endN - S.startN;
            numIterMax -= numIter;
            numIter = 0;
        }
        if (numIter == 0) {
            S.sizeP = 0;
            S.startP = S.endX;
            S.endP = S.startP;
        } else {
            compute_p(S, V, offP);
        }
        {
            Timer t(S, T_COMPX);
            int s[2] = {S.startN, S.startN}, e[2] = {S.endX, S.endX};
            ops->MultiVecAxpby(1.0, ritz, 0.0, V, s, e, ops);
        }
        compute_w(S, V, S.ss_eval.data(), ritz, offW);
        std::swap(offP, offW);
        rayleigh_ritz(S, *nevConv, V);
        for (int i = S.sizeV; i < nevMax + 2 * bs; ++i) S.ss_eval[i] = S.ss_eval[S.sizeV - 1];
        ritz_vec(S, ritz, V);
        ++numIter;
        ++S.indd;
    } while (numIter < numIterMax);
    S.numIter = numIter + (S.numIterMax - numIterMax);
    for (int i = 0; i < S.sizeX; ++i) eval[i] = S.ss_eval[i];
    if (S.orth_method != GCGE_ORTH_MGS) final_rayleigh_quotients(S, ritz, *nevConv, eval, V);
    if (S.printtime) {
        double tot = 0.0;
        for (int k : {T_INITX, T_CHECK, T_COMPP, T_RR, T_RV, T_COMPW, T_COMPX}) tot += S.times[k];
        ops->Printf("|--GCG----------------------------\n");
        ops->Printf("|Total Time = %.2f, Avg Time per Iteration = %.2f\n", tot,
                    tot / std::max(S.numIter, 1));
        ops->Printf("|checkconv   compP   compRR   (rr_matW   dsyexv)   compRV   compW   "
                    "(linsol)   compX   initX\n");
        ops->Printf("|%.2f\t%.2f\t%.2f\t(%.2f\t%.2f)\t%.2f\t%.2f\t(%.2f)\t%.2f\t%.2f\n",
                    S.times[T_CHECK], S.times[T_COMPP], S.times[T_RR], S.times[T_RRW],
                    S.times[T_EIG], S.times[T_RV], S.times[T_COMPW], S.times[T_LINSOL],
                    S.times[T_COMPX], S.times[T_INITX]);
        ops->Printf("|--GCG----------------------------\n");
    }
}

void set_orth(OrthParams &p, int bs, int mr, double zt)
{
    if (bs > 0) p.block_size = bs;
    if (mr >= 0) p.max_reorth = mr;
    if (zt > 0) p.zero_tol = zt;
}

}  // namespace

extern "C" {

void HIPGCG_EigenSolverSetup(double gapMin, int nevInit, int nevMax, int block_size,
                             double tol[2], int numIterMax, int print, int printtime,
                             int user_defined_multi_linear_solver, void **mv_ws[4],
                             double *, int *, OPS *ops)
{
    Solver *S = static_cast<Solver *>(ops->eigen_solver_workspace);
    // reuse a solver previously installed by this function; anything else
    // in the slot belongs to another driver and is left alone
    if (S == nullptr || ops->EigenSolver != gcg) S = new Solver();
    S->gapMin = gapMin;
    S->nevInit = nevInit;
    S->nevMax = nevMax;
    S->block_size = block_size;
    S->tol[0] = tol[0];
    S->tol[1] = tol[1];
    S->numIterMax = numIterMax;
    S->print = print;
    S->printtime = printtime;
    for (int k = 0; k < 4; ++k) S->mv_ws[k] = mv_ws[k];
    S->compRR_min_gap = gapMin;
    S->check_conv_max_num = block_size;
    S->user_ls = user_defined_multi_linear_solver;
    ops->eigen_solver_workspace = S;
    ops->EigenSolver = gcg;
}

void HIPGCG_EigenSolverSetParameters(
    int check_conv_max_num, const char *initX_orth_method, int initX_orth_block_size,
    int initX_orth_max_reorth, double initX_orth_zero_tol, const char *compP_orth_method,
    int compP_orth_block_size, int compP_orth_max_reorth, double compP_orth_zero_tol,
    const char *compW_orth_method, int compW_orth_block_size, int compW_orth_max_reorth,
    double compW_orth_zero_tol, int compW_cg_max_iter, double compW_cg_rate,
    double compW_cg_tol, const char *compW_cg_tol_type, int compW_cg_auto_shift,
    int compRR_min_num, double compRR_min_gap, double compRR_tol, OPS *ops)
{
    Solver &S = *solver_of(ops);
    // only MGS is provided; the method names are accepted as in the
    // reference, where GCGE_Setparameters always passes "mgs"
    (void)initX_orth_method;
    (void)compP_orth_method;
    (void)compW_orth_method;
    if (check_conv_max_num > 0) S.check_conv_max_num = check_conv_max_num;
    set_orth(S.orthX, initX_orth_block_size, initX_orth_max_reorth, initX_orth_zero_tol);
    set_orth(S.orthP, compP_orth_block_size, compP_orth_max_reorth, compP_orth_zero_tol);
    set_orth(S.orthW, compW_orth_block_size, compW_orth_max_reorth, compW_orth_zero_tol);
    if (compW_cg_max_iter > 0) S.cg_max_iter = compW_cg_max_iter;
    if (compW_cg_rate > 0) S.cg_rate = compW_cg_rate;
    if (compW_cg_tol > 0) S.cg_tol = compW_cg_tol;
    if (compW_cg_tol_type != nullptr) S.cg_tol_type = compW_cg_tol_type;
    S.cg_auto_shift = compW_cg_auto_shift;
    if (compRR_min_gap > 0) S.compRR_min_gap = compRR_min_gap;
    if (compRR_min_num > 0) S.compRR_min_num = compRR_min_num;
    if (compRR_tol > 0) S.compRR_tol = compRR_tol;
}

void HIPGCG_Setparameters(double gapMin, int shift, OPS *ops)
{
    HIPGCG_EigenSolverSetParameters(50, "mgs", -1, 2, 2 * DBL_EPSILON, "mgs", -1, 2,
                                    2 * DBL_EPSILON, "mgs", -1, 2, 2 * DBL_EPSILON, 30, 1e-2,
                                    1e-14, "abs", shift, -1, gapMin, 2 * DBL_EPSILON, ops);
}

void HIPGCG_Create(void *A, int nevMax, int block_size, int nevInit, void ***mv_ws,
                   double **dbl_ws, int **int_ws, OPS *ops)
{
    ops->MultiVecCreateByMat(&mv_ws[0], nevMax + 2 * block_size, A, ops);
    for (int k = 1; k < 4; ++k) ops->MultiVecCreateByMat(&mv_ws[k], block_size, A, ops);
    const int sizeV = nevInit + 2 * block_size;
    const int ld = 2 * sizeV * sizeV + 10 * sizeV + (nevMax + 2 * block_size) + nevMax * block_size;
    const int li = 6 * sizeV + 2 * (block_size + 3);
    if (dbl_ws) *dbl_ws = static_cast<double *>(calloc((size_t)ld, sizeof(double)));
    if (int_ws) *int_ws = static_cast<int *>(calloc((size_t)li, sizeof(int)));
}

void HIPGCG_Destroy(void ***mv_ws, double *dbl_ws, int *int_ws, int nevMax, int block_size,
                    OPS *ops)
{
    ops->MultiVecDestroy(&mv_ws[0], nevMax + 2 * block_size, ops);
    for (int k = 1; k < 4; ++k) ops->MultiVecDestroy(&mv_ws[k], block_size, ops);
    free(dbl_ws);
    free(int_ws);
}

void HIPGCG_Free(OPS *ops)
{
    HIPGCG_MultiLinearSolverFree_BlockPCG32(ops);
    if (ops->EigenSolver == gcg && ops->eigen_solver_workspace) {
        delete static_cast<Solver *>(ops->eigen_solver_workspace);
        ops->eigen_solver_workspace = nullptr;
        ops->EigenSolver = nullptr;
    }
}

void HIPGCG_GetStats(OPS *ops, int64_t *stats, double *times)
{
    Solver &S = *solver_of(ops);
    if (stats) {
        stats[0] = S.numIter;
        stats[1] = S.sizeC;
        // FP64 BlockPCG iterations of this driver plus those of a device
        // FP32 BlockPCG installed in the MultiLinearSolver slot
        stats[2] = S.bpcg_iters + (gcge::pcg32_iterations(ops) - S.pcg32_base);
        stats[3] = S.linsolves;
    }
    if (times)
        for (int k = 0; k < 10; ++k) times[k] = S.times[k];
}

void HIPGCG_SetDenseEigensolver(OPS *ops, int which) { solver_of(ops)->eig_backend = which; }

void HIPGCG_SetOrthMethod(OPS *ops, int which) { solver_of(ops)->orth_method = which; }

void HIPGCG_GetOrthStats(OPS *ops, int64_t *cholqr_blocks, int64_t *mgs_fallbacks)
{
    const Solver *S = solver_of(ops);
    *cholqr_blocks = S->cholqr_blocks;
    *mgs_fallbacks = S->cholqr_fallbacks;
}

}  // extern "C"

// Host DenseMatOrth: orthonormalize columns [start, *end) of a small dense
// matrix, the "bqr" option of the P step (app/app_lapack.c:653-699).
//
// Columns [0, start) are projected out twice (classical Gram-Schmidt with
// one re-orthogonalization), then the remaining block is factored by
// Householder QR with column pivoting (Golub & Van Loan, Algorithm 5.4.1);
// the rank is the last pivot with |r_kk| > orth_zero_tol, and the block is
// overwritten by the first `rank` columns of the orthogonal factor.
#include <cmath>
#include <cstring>
#include <vector>

#include "gcge_hip.h"

extern "C" void gcge_host_DenseMatOrth(double *mat, int nrows, int ldm, int start, int *end,
                                       double orth_zero_tol, double *, int, int *)
{
    const int m = nrows;
    int n = *end - start;
    if (n <= 0) return;
    double *A = mat + (size_t)ldm * start;
    if (start > 0) {
        std::vector<double> coef((size_t)start);
        for (int pass = 0; pass < 2; ++pass) {
            for (int j = 0; j < n; ++j) {
                double *a = A + (size_t)ldm * j;
                for (int i = 0; i < start; ++i) {
                    const double *q = mat + (size_t)ldm * i;
                    double s = 0.0;
                    for (int r = 0; r < m; ++r) s += q[r] * a[r];
                    coef[i] = s;
                }
                for (int i = 0; i < start; ++i) {
                    const double *q = mat + (size_t)ldm * i;
                    for (int r = 0; r < m; ++r) a[r] -= coef[i] * q[r];
                }
            }
        }
    }
    const int kmax = m < n ? m : n;
    std::vector<double> tau((size_t)kmax, 0.0), cn((size_t)n), cn0((size_t)n);
    for (int j = 0; j < n; ++j) {
        const double *a = A + (size_t)ldm * j;
        double s = 0.0;
        for (int r = 0; r < m; ++r) s += a[r] * a[r];
        cn[j] = cn0[j] = s;
    }
    for (int k = 0; k < kmax; ++k) {
        int p = k;
        for (int j = k + 1; j < n; ++j)
            if (cn[j] > cn[p]) p = j;
        if (p != k) {
            double *a = A + (size_t)ldm * k, *b = A + (size_t)ldm * p;
            for (int r = 0; r < m; ++r) std::swap(a[r], b[r]);
            std::swap(cn[k], cn[p]);
            std::swap(cn0[k], cn0[p]);
        }
        double *ak = A + (size_t)ldm * k;
        double xn = 0.0;
        for (int r = k; r < m; ++r) xn += ak[r] * ak[r];
        xn = std::sqrt(xn);
        if (xn == 0.0) {
            tau[k] = 0.0;
            continue;
        }
        const double alpha = ak[k] >= 0.0 ? -xn : xn;
        const double v0 = ak[k] - alpha;
        // v = [1, ak[k+1:]/v0], tau = (alpha - ak[k]) / alpha ... stored LAPACK style
        for (int r = k + 1; r < m; ++r) ak[r] /= v0;
        tau[k] = -v0 / alpha;
        ak[k] = alpha;
        for (int j = k + 1; j < n; ++j) {
            double *aj = A + (size_t)ldm * j;
            double s = aj[k];
            for (int r = k + 1; r < m; ++r) s += ak[r] * aj[r];
            s *= tau[k];
            aj[k] -= s;
            for (int r = k + 1; r < m; ++r) aj[r] -= s * ak[r];
            // downdate the remaining column norm; recompute on cancellation
            double t = cn[j] - aj[k] * aj[k];
            if (t <= 1e-10 * cn0[j]) {
                t = 0.0;
                for (int r = k + 1; r < m; ++r) t += aj[r] * aj[r];
                cn0[j] = t;
            }
            cn[j] = t;
        }
    }
    int rank = n;
    for (int c = n - 1; c >= 0; --c) {
        if (c < kmax && std::fabs(A[(size_t)ldm * c + c]) > orth_zero_tol) break;
        rank = c;
    }
    // Q = H_0 ... H_{rank-1} applied to the first `rank` unit vectors.
    const int kk = rank < m ? rank : m;
    std::vector<double> Q((size_t)m * rank, 0.0);
    for (int j = 0; j < rank; ++j) Q[(size_t)m * j + j] = 1.0;
    for (int k = kk - 1; k >= 0; --k) {
        const double *v = A + (size_t)ldm * k;
        for (int j = k; j < rank; ++j) {
            double *q = Q.data() + (size_t)m * j;
            double s = q[k];
            for (int r = k + 1; r < m; ++r) s += v[r] * q[r];
            s *= tau[k];
            q[k] -= s;
            for (int r = k + 1; r < m; ++r) q[r] -= s * v[r];
        }
    }
    for (int j = 0; j < rank; ++j)
        memcpy(A + (size_t)ldm * j, Q.data() + (size_t)m * j, sizeof(double) * m);
    *end = start + rank;
}

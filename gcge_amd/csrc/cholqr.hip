// The small dense step of a block Cholesky QR pass on the device: from the
// B-Gram matrix G = X^T B X of a block (written by the Gram kernel, lower
// triangle), the coefficients C = D^{-1} R^{-1} of X <- X C, where
// D = diag(G)^{1/2} and D^{-1} G D^{-1} = R^T R.  One workgroup; the
// factorization is right-looking Cholesky of L = R^T in LDS (orders up to
// 128) or in a global work array (up to 1024), and R^{-1} = (L^{-1})^T by
// forward substitution, one column of L^{-1} per thread.
//
// The tests are those of the host factorization in the GCG driver: a
// column whose B-norm is below zero_tol, a non-positive largest scaled
// diagonal entry, or a pivot not above rel_tol times it marks the block
// numerically dependent.  Then flag[0] = 1 and C is the identity, so the
// linear combination that follows leaves X as it was and the caller falls
// back to column-by-column Gram-Schmidt from the same state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "app_hip_internal.h"
#include "kernels.h"

namespace gcge {
namespace {

constexpr int kCqThreads = 1024;            // 16 waves
constexpr int kCqWaves = kCqThreads / 64;
constexpr int kCqLds = 128;                 // orders factored in LDS
constexpr int kCqMax = 1024;                // one column of L^{-1} per thread

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// A: column-major n x n (lda = n); L in the lower triangle, its diagonal
// in ldiag, and Z = L^{-1} below the diagonal stored transposed in the
// strict upper triangle (Z[i][j], i > j, at A[j + n i]).
template <int kN, bool kLds>
__global__ __launch_bounds__(kCqThreads) void k_cholqr_coef(
    int n, const double *__restrict__ G, int ldg, double zero_tol, double rel_tol,
    double *__restrict__ work, double *__restrict__ coef, int colmajor,
    double *__restrict__ flag)
{
    __shared__ double lds_a[kLds ? kN * kN : 1];
    __shared__ double dinv[kN], ldiag[kN], red[kCqWaves];
    __shared__ int s_bad;
    double *A = kLds ? lds_a : work;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;

    if (t == 0) s_bad = 0;
    __syncthreads();
    // D^{-1}; a column of B-norm below zero_tol fails the block
    for (int i = t; i < n; i += kCqThreads) {
        const double d = sqrt(fmax(G[i + (int64_t)ldg * i], 0.0));
        const bool ok = d >= zero_tol;
        dinv[i] = ok ? 1.0 / d : 0.0;
        if (!ok) s_bad = 1;
    }
    __syncthreads();
    bool bad = s_bad != 0;

    // A = D^{-1} G D^{-1} (lower triangle) and its largest diagonal entry
    if (!bad) {
        for (int j = wv; j < n; j += kCqWaves)
            for (int i = j + lane; i < n; i += 64)
                A[i + n * j] = G[i + (int64_t)ldg * j] * (dinv[i] * dinv[j]);
        double m = 0.0;
        for (int i = t; i < n; i += kCqThreads)
            m = fmax(m, G[i + (int64_t)ldg * i] * (dinv[i] * dinv[i]));
        m = wave_max(m);
        if (lane == 0) red[wv] = m;
    }
    __syncthreads();
    double dmax = 0.0;
    if (!bad) {
        for (int w = 0; w < kCqWaves; ++w) dmax = fmax(dmax, red[w]);
        bad = !(dmax > 0.0);
    }

    // right-looking Cholesky: every thread reads the same pivot after the
    // barrier that ends the previous step, so all leave the loop together
    for (int k = 0; k < n && !bad; ++k) {
        const double s = A[k + n * k];
        if (!(s > rel_tol * dmax)) {
            bad = true;
            break;
        }
        const double lkk = sqrt(s);
        if (t == 0) ldiag[k] = lkk;
        for (int i = k + 1 + t; i < n; i += kCqThreads) A[i + n * k] /= lkk;
        __syncthreads();
        for (int j = k + 1 + wv; j < n; j += kCqWaves) {
            const double ljk = A[j + n * k];
            for (int i = j + lane; i < n; i += 64) A[i + n * j] -= A[i + n * k] * ljk;
        }
        __syncthreads();
    }

    // Z = L^{-1}: thread j computes column j top-down,
    // Z[i][j] = -(L[i][j] / L[j][j] + sum_{j<k<i} L[i][k] Z[k][j]) / L[i][i];
    // it writes only row j of the upper triangle and reads only L and that row
    if (!bad && t < n) {
        const int j = t;
        const double zjj = 1.0 / ldiag[j];
        for (int i = j + 1; i < n; ++i) {
            double acc = A[i + n * j] * zjj;
            for (int k = j + 1; k < i; ++k) acc += A[i + n * k] * A[j + n * k];
            A[j + n * i] = -acc / ldiag[i];
        }
    }
    __syncthreads();

    // C[i][k] = dinv[i] (R^{-1})[i][k], (R^{-1})[i][k] = Z[k][i] (k >= i);
    // the identity for a dependent block
    for (int k = wv; k < n; k += kCqWaves)
        for (int i = lane; i < n; i += 64) {
            double v;
            if (bad)
                v = (i == k) ? 1.0 : 0.0;
            else if (i > k)
                v = 0.0;
            else if (i == k)
                v = (1.0 / ldiag[i]) * dinv[i];
            else
                v = A[i + n * k] * dinv[i];
            coef[colmajor ? i + (int64_t)n * k : (int64_t)i * n + k] = v;
        }
    if (t == 0) flag[0] = bad ? 1.0 : 0.0;
}

}  // namespace

int64_t cholqr_coef_work(int n) { return n > kCqLds ? (int64_t)n * n : 0; }

bool cholqr_coef(int n, const double *G, int ldg, double zero_tol, double rel_tol,
                 double *work, double *coef, bool colmajor, double *flag, hipStream_t s)
{
    if (n < 1 || n > kCqMax || ldg < n) return false;
    if (n <= kCqLds)
        k_cholqr_coef<kCqLds, true><<<1, kCqThreads, 0, s>>>(
            n, G, ldg, zero_tol, rel_tol, nullptr, coef, colmajor ? 1 : 0, flag);
    else
        k_cholqr_coef<kCqMax, false><<<1, kCqThreads, 0, s>>>(
            n, G, ldg, zero_tol, rel_tol, work, coef, colmajor ? 1 : 0, flag);
    HIPCHECK(hipGetLastError());
    return true;
}

}  // namespace gcge

// The small dense step of a block Cholesky QR pass on the device: from the
// B-Gram matrix G = X^T B X of a block (written by the Gram kernel, lower
// triangle), the coefficients C = D^{-1} R^{-1} of X <- X C, where
// D = diag(G)^{1/2} and D^{-1} G D^{-1} = R^T R.  One workgroup; the
// factorization is right-looking Cholesky of L = R^T in LDS (orders up to
// 128) or in a global work array (up to 1024), one barrier per column, and
// R^{-1} = (L^{-1})^T by column-oriented forward substitution, one column
// of L^{-1} per wavefront held in registers (the pivot entry goes to every
// lane by readlane), so the substitution needs no barrier and no LDS
// traffic beyond L itself.
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
constexpr int kCqLds = 128;                 // orders factored in LDS
constexpr int kCqMax = 1024;

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// v of lane l (l uniform across the wavefront)
__device__ inline double readlane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// A: column-major n x n (lda = n), the scaled Gram matrix and then L in
// its lower triangle; rdiag[k] = 1 / L[k][k].
template <int kN, bool kLds, int kThreads>
__global__ __launch_bounds__(kThreads) void k_cholqr_coef(
    int n, const double *__restrict__ G, int ldg, double zero_tol, double rel_tol,
    double *__restrict__ work, double *__restrict__ coef, int colmajor,
    double *__restrict__ flag)
{
    constexpr int kWaves = kThreads / 64, kPer = kN / 64;
    __shared__ double lds_a[kLds ? kN * kN : 1];
    __shared__ double dinv[kN], rdiag[kN], red[kWaves];
    __shared__ int s_bad;
    double *A = kLds ? lds_a : work;
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // uniform per wavefront

    if (t == 0) s_bad = 0;
    __syncthreads();
    // D^{-1}; a column of B-norm below zero_tol fails the block
    for (int i = t; i < n; i += kThreads) {
        const double d = sqrt(fmax(G[i + (int64_t)ldg * i], 0.0));
        const bool ok = d >= zero_tol;
        dinv[i] = ok ? 1.0 / d : 0.0;
        if (!ok) s_bad = 1;
    }
    __syncthreads();
    bool bad = s_bad != 0;

    // A = D^{-1} G D^{-1} (lower triangle) and its largest diagonal entry
    if (!bad) {
        for (int j = wv; j < n; j += kWaves)
            for (int i = j + lane; i < n; i += 64)
                A[i + n * j] = G[i + (int64_t)ldg * j] * (dinv[i] * dinv[j]);
        double m = 0.0;
        for (int i = t; i < n; i += kThreads)
            m = fmax(m, G[i + (int64_t)ldg * i] * (dinv[i] * dinv[i]));
        m = wave_max(m);
        if (lane == 0) red[wv] = m;
    }
    __syncthreads();
    double dmax = 0.0;
    if (!bad) {
        for (int w = 0; w < kWaves; ++w) dmax = fmax(dmax, red[w]);
        bad = !(dmax > 0.0);
    }

    // Right-looking Cholesky, one barrier per column.  Step k reads the
    // pivot and column k as step k-1 left them (every thread reads the same
    // pivot, so all leave the loop together), updates columns k+1.. with
    // L[i][k] = A[i][k] / sqrt(A[k][k]) formed on the fly, and stores
    // column k-1 of L, which no thread reads or writes during step k.
    double rprev = 0.0;
    for (int k = 0; k < n && !bad; ++k) {
        const double s = A[k + n * k];
        if (!(s > rel_tol * dmax)) {
            bad = true;
            break;
        }
        const double rl = 1.0 / sqrt(s);
        if (t == 0) rdiag[k] = rl;
        if (k > 0)
            for (int i = k + t; i < n; i += kThreads) A[i + n * (k - 1)] *= rprev;
        for (int j = k + 1 + wv; j < n; j += kWaves) {
            const double ljk = A[j + n * k] * rl;
            for (int i = j + lane; i < n; i += 64) A[i + n * j] -= (A[i + n * k] * rl) * ljk;
        }
        rprev = rl;
        __syncthreads();
    }

    if (bad) {
        for (int k = wv; k < n; k += kWaves)
            for (int i = lane; i < n; i += 64)
                coef[colmajor ? i + (int64_t)n * k : (int64_t)i * n + k] = (i == k) ? 1.0 : 0.0;
    } else {
        // Column j of Z = L^{-1} per wavefront, row i = lane + 64 r in z[r]:
        // for k = j..n-1, z_k /= L[k][k], then z_i -= L[i][k] z_k (i > k).
        // Z[i][j] is (R^{-1})[j][i], so the wavefront writes row j of
        // C = D^{-1} R^{-1}: C[j][i] = dinv[j] Z[i][j] (i >= j), else 0.
        for (int j = wv; j < n; j += kWaves) {
            double z[kPer];
#pragma unroll
            for (int r = 0; r < kPer; ++r) z[r] = (lane + 64 * r == j) ? 1.0 : 0.0;
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                if (64 * r >= n) break;
                for (int kk = 0; kk < 64; ++kk) {
                    const int k = 64 * r + kk;
                    if (k >= n) break;
                    if (k < j) continue;
                    const double zk = readlane(z[r], kk) * rdiag[k];
                    if (lane == kk) z[r] = zk;
#pragma unroll
                    for (int q = r; q < kPer; ++q) {
                        const int i = lane + 64 * q;
                        if (i > k && i < n) z[q] -= A[i + n * k] * zk;
                    }
                }
            }
            const double dj = dinv[j];
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const int i = lane + 64 * r;
                if (i < n)
                    coef[colmajor ? j + (int64_t)n * i : (int64_t)j * n + i] =
                        i >= j ? z[r] * dj : 0.0;
            }
        }
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
        k_cholqr_coef<kCqLds, true, kCqThreads><<<1, kCqThreads, 0, s>>>(
            n, G, ldg, zero_tol, rel_tol, nullptr, coef, colmajor ? 1 : 0, flag);
    else
        k_cholqr_coef<kCqMax, false, kCqThreads><<<1, kCqThreads, 0, s>>>(
            n, G, ldg, zero_tol, rel_tol, work, coef, colmajor ? 1 : 0, flag);
    HIPCHECK(hipGetLastError());
    return true;
}

}  // namespace gcge

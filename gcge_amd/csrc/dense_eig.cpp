// Dense symmetric eigensolvers for the projected (Rayleigh-Ritz) problem.
//
// eigh_host: Householder tridiagonalization and implicit symmetric QR with
// Wilkinson shifts (Golub & Van Loan, Alg. 8.3.1 and 8.3.3), about 9 n^3
// flops with the eigenvectors, for the small projected matrices of small
// problems (no launches, no device round trip).
// eigh_device: rocSOLVER dsyevd (Householder tridiagonalization + divide
// and conquer) on the device for the larger projected matrices.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "app_hip_internal.h"
#include "dense_eig.h"

namespace gcge {

int eigh_host(int n, double *A, int lda, double *w, double *Z, int ldz)
{
    if (n <= 0) return 0;
    const size_t nn = (size_t)n * n;
    // a: full symmetric copy of the upper triangle, column-major
    std::vector<double> a(nn), T(nn, 0.0), Q(nn, 0.0), beta((size_t)n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int r = std::min(i, j), c = std::max(i, j);
            a[(size_t)n * j + i] = A[(size_t)lda * c + r];
        }
    auto at = [&](int i, int j) -> double & { return a[(size_t)n * j + i]; };
    std::vector<double> d((size_t)n), e((size_t)n, 0.0), pv((size_t)n);
    // 1. Householder tridiagonalization (Golub & Van Loan, Alg. 8.3.1):
    // column k's reflector I - beta v v^T (v stored in a[k+1:, k]) zeroes
    // a[k+2:, k]; the trailing block is updated as S - v w^T - w v^T with
    // p = beta S v, w = p - (beta p^T v / 2) v.
    for (int k = 0; k + 2 < n; ++k) {
        const int L = n - k - 1;
        double *x = &at(k + 1, k);
        double nrm2 = 0.0;
        for (int i = 0; i < L; ++i) nrm2 += x[i] * x[i];
        double alpha = std::sqrt(nrm2);
        if (alpha == 0.0) {
            e[k] = 0.0;
            beta[k] = 0.0;
            continue;
        }
        if (x[0] > 0) alpha = -alpha;
        const double vtv = 2.0 * (nrm2 - x[0] * alpha);
        x[0] -= alpha;
        const double b = 2.0 / vtv;
        beta[k] = b;
        e[k] = alpha;
        // p = beta S v, column by column (contiguous axpys)
        std::fill(pv.begin(), pv.begin() + L, 0.0);
        for (int jj = 0; jj < L; ++jj) {
            const double vj = x[jj];
            const double *col = &at(k + 1, k + 1 + jj);
            for (int i = 0; i < L; ++i) pv[i] += col[i] * vj;
        }
        double ptv = 0.0;
        for (int i = 0; i < L; ++i) {
            pv[i] *= b;
            ptv += pv[i] * x[i];
        }
        const double K = 0.5 * b * ptv;
        for (int i = 0; i < L; ++i) pv[i] -= K * x[i];
        for (int jj = 0; jj < L; ++jj)
            for (int i = 0; i < L; ++i)
                at(k + 1 + i, k + 1 + jj) -= x[i] * pv[jj] + pv[i] * x[jj];
    }
    for (int k = 0; k < n; ++k) d[k] = at(k, k);
    if (n >= 2) e[n - 2] = at(n - 1, n - 2);
    // Q = H_0 H_1 ... H_{n-3}, accumulated backward
    for (int k = 0; k < n; ++k) Q[(size_t)n * k + k] = 1.0;
    for (int k = n - 3; k >= 0; --k) {
        if (beta[k] == 0.0) continue;
        const int L = n - k - 1;
        const double *v = &at(k + 1, k);
        for (int jj = k + 1; jj < n; ++jj) {
            double *q = &Q[(size_t)n * jj + k + 1];
            double sacc = 0.0;
            for (int i = 0; i < L; ++i) sacc += v[i] * q[i];
            sacc *= beta[k];
            for (int i = 0; i < L; ++i) q[i] -= sacc * v[i];
        }
    }
    // 2. Implicit symmetric QR with Wilkinson shifts on the tridiagonal
    // (Golub & Van Loan, Alg. 8.3.3): T is kept dense, each Givens similarity
    // touches its 4 x 4 band window, and the rotations accumulate into Q.
    auto t = [&](int i, int j) -> double & { return T[(size_t)n * j + i]; };
    for (int k = 0; k < n; ++k) {
        t(k, k) = d[k];
        if (k + 1 < n) t(k + 1, k) = t(k, k + 1) = e[k];
    }
    const double eps = 2.220446049250313e-16;
    auto rot = [&](int k, int lo, int hi, double c, double s) {
        for (int jj = lo; jj <= hi; ++jj) {
            const double t1 = t(k, jj), t2 = t(k + 1, jj);
            t(k, jj) = c * t1 - s * t2;
            t(k + 1, jj) = s * t1 + c * t2;
        }
        for (int ii = lo; ii <= hi; ++ii) {
            const double t1 = t(ii, k), t2 = t(ii, k + 1);
            t(ii, k) = c * t1 - s * t2;
            t(ii, k + 1) = s * t1 + c * t2;
        }
        double *z1 = &Q[(size_t)n * k], *z2 = &Q[(size_t)n * (k + 1)];
        for (int ii = 0; ii < n; ++ii) {
            const double a1 = z1[ii], a2 = z2[ii];
            z1[ii] = c * a1 - s * a2;
            z2[ii] = s * a1 + c * a2;
        }
    };
    int steps = 0;
    for (;;) {
        for (int i = 0; i + 1 < n; ++i)
            if (std::fabs(t(i + 1, i)) <= eps * (std::fabs(t(i, i)) + std::fabs(t(i + 1, i + 1))))
                t(i + 1, i) = t(i, i + 1) = 0.0;
        int m = n - 1;
        while (m > 0 && t(m, m - 1) == 0.0) --m;
        if (m == 0) break;
        int p = m - 1;
        while (p > 0 && t(p, p - 1) != 0.0) --p;
        if (++steps > 30 * n) return 1;  // no convergence
        const double dd = 0.5 * (t(m - 1, m - 1) - t(m, m));
        const double em = t(m, m - 1);
        const double den = dd + std::copysign(std::hypot(dd, em), dd);
        const double mu = t(m, m) - (den != 0.0 ? em * em / den : 0.0);
        for (int k = p; k < m; ++k) {
            double x, z;
            if (k == p) {
                x = t(p, p) - mu;
                z = t(p + 1, p);
            } else {
                x = t(k, k - 1);
                z = t(k + 1, k - 1);
            }
            const double r = std::hypot(x, z);
            const double c = r != 0.0 ? x / r : 1.0, s = r != 0.0 ? -z / r : 0.0;
            rot(k, std::max(p, k - 1), std::min(m, k + 2), c, s);
            if (k > p) t(k + 1, k - 1) = t(k - 1, k + 1) = 0.0;
        }
    }
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return t(x, x) < t(y, y); });
    for (int k = 0; k < n; ++k) {
        w[k] = t(order[k], order[k]);
        memcpy(Z + (size_t)ldz * k, Q.data() + (size_t)n * order[k], sizeof(double) * n);
    }
    return 0;
}

namespace {
// per thread (held by the thread's context): a rocBLAS handle on the
// thread's stream, the device copies of the projected matrix and pinned
// host staging for them (contiguous copies at DMA speed instead of 2D
// copies from pageable memory)
struct DevEig {
    rocblas_handle h = nullptr;
    double *dA = nullptr, *dW = nullptr, *dE = nullptr;
    rocblas_int *dinfo = nullptr;
    double *hA = nullptr, *hW = nullptr;
    rocblas_int *hinfo = nullptr;
    int cap = 0;
};

void free_buffers(DevEig &d)
{
    if (d.dA == nullptr) return;
    dev_free(d.dA);
    dev_free(d.dW);
    dev_free(d.dE);
    dev_free(d.dinfo);
    pinned_free(d.hA);
    pinned_free(d.hW);
    pinned_free(d.hinfo);
    d.dA = d.dW = d.dE = d.hA = d.hW = nullptr;
    d.dinfo = d.hinfo = nullptr;
    d.cap = 0;
}
DevEig &deveig()
{
    Ctx &c = ctx();
    if (c.dev_eig == nullptr) {
        DevEig *d = new DevEig();
        if (rocblas_create_handle(&d->h) != rocblas_status_success)
            fail("rocblas_create_handle failed", __FILE__, __LINE__);
        rocblas_set_stream(d->h, c.stream);
        c.dev_eig = d;
    }
    return *static_cast<DevEig *>(c.dev_eig);
}
}  // namespace

void release_dev_eig(void *p)
{
    DevEig *d = static_cast<DevEig *>(p);
    free_buffers(*d);
    rocblas_destroy_handle(d->h);
    delete d;
}

void gemm_nn(int64_t m, int n, int k, const double *A, int64_t lda, const double *B, int ldb,
             double beta, double *C, int64_t ldc)
{
    if (m > INT32_MAX || lda > INT32_MAX || ldc > INT32_MAX)
        fail("gemm_nn: dimensions past the library's 32-bit range", __FILE__, __LINE__);
    DevEig &d = deveig();
    const double one = 1.0;
    const rocblas_status st =
        rocblas_dgemm(d.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, n, k,
                      &one, A, (rocblas_int)lda, B, ldb, &beta, C, (rocblas_int)ldc);
    if (st != rocblas_status_success) fail("rocblas_dgemm failed", __FILE__, __LINE__);
}

void gemm_small(bool transa, int m, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc)
{
    DevEig &d = deveig();
    const rocblas_status st = rocblas_dgemm(
        d.h, transa ? rocblas_operation_transpose : rocblas_operation_none,
        rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc);
    if (st != rocblas_status_success) fail("rocblas_dgemm failed", __FILE__, __LINE__);
}

void transpose(int m, int n, const double *A, int lda, double *C, int ldc)
{
    DevEig &d = deveig();
    const double one = 1.0, zero = 0.0;
    const rocblas_status st =
        rocblas_dgeam(d.h, rocblas_operation_transpose, rocblas_operation_transpose, n, m, &one,
                      A, lda, &zero, A, lda, C, ldc);
    if (st != rocblas_status_success) fail("rocblas_dgeam failed", __FILE__, __LINE__);
}

int eigh_device(int n, double *A, int lda, double *w, double *Z, int ldz)
{
    if (n <= 0) return 0;
    DevEig &d = deveig();
    Ctx &c = ctx();
    if (n > d.cap) {
        c.sync();
        free_buffers(d);
        d.cap = n + 64;
        const size_t sq = sizeof(double) * (size_t)d.cap * d.cap;
        dev_alloc(&d.dA, sq);
        dev_alloc(&d.dW, sizeof(double) * d.cap);
        dev_alloc(&d.dE, sizeof(double) * d.cap);
        dev_alloc(&d.dinfo, sizeof(rocblas_int));
        pinned_alloc(&d.hA, sq);
        pinned_alloc(&d.hW, sizeof(double) * d.cap);
        pinned_alloc(&d.hinfo, sizeof(rocblas_int));
    }
    const size_t col = sizeof(double) * (size_t)n;
    for (int j = 0; j < n; ++j) memcpy(d.hA + (size_t)n * j, A + (size_t)lda * j, col);
    HIPCHECK(hipMemcpyAsync(d.dA, d.hA, col * n, hipMemcpyHostToDevice, c.stream));
    rocblas_status st = rocsolver_dsyevd(d.h, rocblas_evect_original, rocblas_fill_upper, n,
                                         d.dA, n, d.dW, d.dE, d.dinfo);
    if (st != rocblas_status_success) {
        c.sync();  // the staging buffer stays in use until the copy is done
        return -1000 - (int)st;
    }
    HIPCHECK(hipMemcpyAsync(d.hinfo, d.dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                            c.stream));
    HIPCHECK(hipMemcpyAsync(d.hW, d.dW, col, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipMemcpyAsync(d.hA, d.dA, col * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    memcpy(w, d.hW, col);
    for (int j = 0; j < n; ++j) memcpy(Z + (size_t)ldz * j, d.hA + (size_t)n * j, col);
    return (int)*d.hinfo;
}

}  // namespace gcge

extern "C" int gcge_host_dense_eigh(int n, const double *A, int lda, double *w, double *Z,
                                    int ldz)
{
    std::vector<double> a((size_t)n * (n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) memcpy(&a[(size_t)n * j], A + (size_t)lda * j, sizeof(double) * n);
    return gcge::eigh_host(n, a.data(), n, w, Z, ldz);
}

// Dense symmetric eigensolvers for the projected (Rayleigh-Ritz) problem.
//
// eigh_host: cyclic Jacobi with a threshold strategy (Golub & Van Loan,
// Sec. 8.5), accurate to a few ulps relative to ||A|| and suited to the
// small projected matrices of small problems.
// eigh_device: rocSOLVER dsyevd (Householder tridiagonalization + divide
// and conquer) on the device for the larger projected matrices.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
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
    std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int r = std::min(i, j), c = std::max(i, j);
            a[(size_t)n * j + i] = A[(size_t)lda * c + r];  // upper triangle
        }
    for (int i = 0; i < n; ++i) v[(size_t)n * i + i] = 1.0;
    auto at = [&](int i, int j) -> double & { return a[(size_t)n * j + i]; };
    double fro = 0.0;
    for (double x : a) fro += x * x;
    fro = std::sqrt(fro);
    const double eps = 1e-17;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off += at(i, j) * at(i, j);
        off = std::sqrt(2.0 * off);
        if (off <= eps * fro || fro == 0.0) break;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double apq = at(p, q);
                if (std::fabs(apq) <= 1e-300) continue;
                const double app = at(p, p), aqq = at(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                // A <- J^T A J on rows/columns p, q
                for (int k = 0; k < n; ++k) {
                    const double akp = at(k, p), akq = at(k, q);
                    at(k, p) = c * akp - s * akq;
                    at(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = at(p, k), aqk = at(q, k);
                    at(p, k) = c * apk - s * aqk;
                    at(q, k) = s * apk + c * aqk;
                }
                at(p, q) = at(q, p) = 0.0;
                double *vp = v.data() + (size_t)n * p, *vq = v.data() + (size_t)n * q;
                for (int k = 0; k < n; ++k) {
                    const double x = vp[k], y = vq[k];
                    vp[k] = c * x - s * y;
                    vq[k] = s * x + c * y;
                }
            }
        }
    }
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return at(x, x) < at(y, y); });
    for (int k = 0; k < n; ++k) {
        w[k] = at(order[k], order[k]);
        memcpy(Z + (size_t)ldz * k, v.data() + (size_t)n * order[k], sizeof(double) * n);
    }
    return 0;
}

namespace {
// per thread (held by the thread's context): a rocBLAS handle on the
// thread's stream and the device copies of the projected matrix
struct DevEig {
    rocblas_handle h = nullptr;
    double *dA = nullptr, *dW = nullptr, *dE = nullptr;
    rocblas_int *dinfo = nullptr;
    int cap = 0;
};
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
    if (d->dA) {
        HIPCHECK(hipFree(d->dA));
        HIPCHECK(hipFree(d->dW));
        HIPCHECK(hipFree(d->dE));
        HIPCHECK(hipFree(d->dinfo));
    }
    rocblas_destroy_handle(d->h);
    delete d;
}

int eigh_device(int n, double *A, int lda, double *w, double *Z, int ldz)
{
    if (n <= 0) return 0;
    DevEig &d = deveig();
    Ctx &c = ctx();
    if (n > d.cap) {
        c.sync();
        if (d.dA) {
            HIPCHECK(hipFree(d.dA));
            HIPCHECK(hipFree(d.dW));
            HIPCHECK(hipFree(d.dE));
            HIPCHECK(hipFree(d.dinfo));
        }
        d.cap = n + 64;
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&d.dA), sizeof(double) * d.cap * d.cap));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&d.dW), sizeof(double) * d.cap));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&d.dE), sizeof(double) * d.cap));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&d.dinfo), sizeof(rocblas_int)));
    }
    HIPCHECK(hipMemcpy2DAsync(d.dA, sizeof(double) * n, A, sizeof(double) * lda,
                              sizeof(double) * n, n, hipMemcpyHostToDevice, c.stream));
    rocblas_status st = rocsolver_dsyevd(d.h, rocblas_evect_original, rocblas_fill_upper, n,
                                         d.dA, n, d.dW, d.dE, d.dinfo);
    if (st != rocblas_status_success) return -1000 - (int)st;
    rocblas_int info = 0;
    HIPCHECK(hipMemcpyAsync(&info, d.dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipMemcpyAsync(w, d.dW, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipMemcpy2DAsync(Z, sizeof(double) * ldz, d.dA, sizeof(double) * n,
                              sizeof(double) * n, n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return (int)info;
}

}  // namespace gcge

// Host small-dense operation table.
//
// The GCG driver keeps every small matrix (projected matrix, Ritz
// coefficients, P coefficients) on the host and orthogonalizes the P
// coefficients through ops->lapack_ops (src/ops_eig_sol_gcg.c:386-419);
// OPS_Setup copies DenseMatQtAP/DenseMatOrth from it (src/ops.c:97-106).
// This table provides those host operations over the LAPACKVEC layout
// (gcge_dense_mv: column-major, ld >= nrows), with the slot semantics of
// app/app_lapack.c:24-699.  Sizes here are at most a few hundred, so the
// arithmetic is plain cache-ordered loops; the larger products spread their
// output columns over a few threads (the GPU waits for them).
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cassert>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <thread>
#include <vector>

#include "gcge_ops.h"
#include "gcge_hip.h"

namespace {

// Column-parallel loops: f(j) for j in [0, n), on up to kMaxThreads threads
// (fewer when the process may run on fewer CPUs) once a product has
// kParallelWork multiply-adds: the projected matrices of a 100-pair solve
// (order 460 times 120 columns, 25 M) are split, those of small problems
// (a few million, solved several at a time by the service's workers) are
// not.  Each j is computed by one thread exactly as the serial loop
// computes it, so the results do not depend on the thread count.
constexpr int kMaxThreads = 8;
constexpr double kParallelWork = 8e6;

int usable_cpus()
{
    static const int n = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof(set), &set) != 0) return 1;
        return std::max(1, CPU_COUNT(&set));
    }();
    return n;
}

template <class F>
void for_columns(int n, double work, F &&f)
{
    int nt = work < kParallelWork ? 1 : std::min({kMaxThreads, usable_cpus(), n});
    if (nt <= 1) {
        for (int j = 0; j < n; ++j) f(j);
        return;
    }
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (int j; (j = next.fetch_add(1)) < n;) f(j);
    };
    std::vector<std::thread> pool;
    pool.reserve((size_t)nt - 1);
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
}

using mv = gcge_dense_mv;

inline mv *as_mv(void *p) { return static_cast<mv *>(p); }
inline mv *as_mv(void **p) { return reinterpret_cast<mv *>(p); }

double *col(mv *x, int c) { return x->data + (size_t)x->ldd * (size_t)c; }

// a . b over k entries with four independent partial sums, so the sum is
// not one dependency chain (a single accumulator runs at the latency of
// one FMA per element: the compiler may not reassociate it)
double dot(const double *a, const double *b, int k)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int r = 0;
    for (; r + 4 <= k; r += 4) {
        s0 += a[r] * b[r];
        s1 += a[r + 1] * b[r + 1];
        s2 += a[r + 2] * b[r + 2];
        s3 += a[r + 3] * b[r + 3];
    }
    for (; r < k; ++r) s0 += a[r] * b[r];
    return (s0 + s1) + (s2 + s3);
}

// C[0:mc, 0:nc] (ldC) = alpha * Q^T P + beta * C, Q: k x mc (ldQ), P: k x nc.
// Four columns of Q share each pass over a column of P, two partial sums
// each.
void gemm_tn(int mc, int nc, int k, double alpha, const double *Q, int ldQ,
             const double *P, int ldP, double beta, double *C, int ldC)
{
    auto put = [&](int i, int j, double v) {
        double *c = C + (size_t)ldC * j + i;
        *c = (beta == 0.0) ? alpha * v : alpha * v + beta * *c;
    };
    for_columns(nc, (double)mc * nc * k, [&](int j) {
        const double *p = P + (size_t)ldP * j;
        int i = 0;
        for (; i + 4 <= mc; i += 4) {
            const double *q0 = Q + (size_t)ldQ * i, *q1 = q0 + ldQ, *q2 = q1 + ldQ,
                         *q3 = q2 + ldQ;
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0, b3 = 0;
            int r = 0;
            for (; r + 2 <= k; r += 2) {
                const double p0 = p[r], p1 = p[r + 1];
                a0 += q0[r] * p0; b0 += q0[r + 1] * p1;
                a1 += q1[r] * p0; b1 += q1[r + 1] * p1;
                a2 += q2[r] * p0; b2 += q2[r + 1] * p1;
                a3 += q3[r] * p0; b3 += q3[r + 1] * p1;
            }
            if (r < k) {
                a0 += q0[r] * p[r]; a1 += q1[r] * p[r];
                a2 += q2[r] * p[r]; a3 += q3[r] * p[r];
            }
            put(i, j, a0 + b0);
            put(i + 1, j, a1 + b1);
            put(i + 2, j, a2 + b2);
            put(i + 3, j, a3 + b3);
        }
        for (; i < mc; ++i) put(i, j, dot(Q + (size_t)ldQ * i, p, k));
    });
}

// C (m x n, ldC) = A (m x k, ldA) * B (k x n, ldB), overwriting C.
void gemm_nn(int m, int n, int k, const double *A, int ldA, const double *B, int ldB,
             double *C, int ldC)
{
    for_columns(n, (double)m * n * k, [&](int j) {
        double *c = C + (size_t)ldC * j;
        for (int i = 0; i < m; ++i) c[i] = 0.0;
        for (int p = 0; p < k; ++p) {
            const double b = B[(size_t)ldB * j + p];
            if (b == 0.0) continue;
            const double *a = A + (size_t)ldA * p;
            for (int i = 0; i < m; ++i) c[i] += a[i] * b;
        }
    });
}

}  // namespace

extern "C" {

// matC = alpha * matQ^T op(matA) matP + beta * matC   (app/app_lapack.c:24-227)
// ntluA: 'L'/'U' -> matA symmetric (that triangle is read), 'N'/'T' -> general.
// nsdC: 'N' full, 'S' symmetric (lower computed, mirrored), 'D' diagonal.
void gcge_host_DenseMatQtAP(char ntluA, char nsdC, int nrowsA, int ncolsA,
                            int nrowsC, int ncolsC, double alpha, double *matQ,
                            int ldQ, double *matA, int ldA, double *matP, int ldP,
                            double beta, double *matC, int ldC, double *dbl_ws)
{
    if (nrowsC == 0 || ncolsC == 0) return;
    if (nrowsA == 0 && ncolsA == 0) {
        if (nsdC == 'D') {
            for (int c = 0; c < ncolsC; ++c) matC[(size_t)ldC * c] = 0.0;
        } else {
            for (int c = 0; c < ncolsC; ++c)
                memset(matC + (size_t)ldC * c, 0, sizeof(double) * nrowsC);
        }
        return;
    }
    if (matA == nullptr) {
        const int k = nrowsA;
        if (nsdC == 'D') {
            for (int c = 0; c < ncolsC; ++c) {
                const double *q = matQ + (size_t)ldQ * c, *p = matP + (size_t)ldP * c;
                const double s = dot(q, p, k);
                double *d = matC + (size_t)ldC * c;
                if (beta == 1.0) {
                    if (alpha != 0.0) *d += alpha * s;
                } else if (beta == 0.0) {
                    *d = alpha * s;
                } else {
                    *d = alpha * s + beta * *d;
                }
            }
        } else if (nsdC == 'S') {
            // column c writes rows c.. of column c, then row c of the
            // columns after it: no entry is written by two columns
            for_columns(ncolsC, 0.5 * (double)ncolsC * ncolsC * k, [&](int c) {
                const double *p = matP + (size_t)ldP * c;
                for (int r = c; r < ncolsC; ++r) {
                    const double s = dot(matQ + (size_t)ldQ * r, p, k);
                    double *d = matC + (size_t)ldC * c + r;
                    *d = (beta == 0.0) ? alpha * s : alpha * s + beta * *d;
                }
                for (int r = c + 1; r < ncolsC; ++r)
                    matC[(size_t)ldC * r + c] = matC[(size_t)ldC * c + r];
            });
        } else {
            gemm_tn(nrowsC, ncolsC, k, alpha, matQ, ldQ, matP, ldP, beta, matC, ldC);
        }
        return;
    }
    // W = op(A) P into dbl_ws (nrowsA x ncolsC), then the A == NULL case.
    std::vector<double> own;
    double *W = dbl_ws;
    if (W == nullptr) {
        own.resize((size_t)nrowsA * ncolsC);
        W = own.data();
    }
    const int ldW = nrowsA;
    if (ntluA == 'L' || ntluA == 'U') {
        // symmetric A, only the `ntluA` triangle is valid
        const int n = nrowsA;
        for_columns(ncolsC, (double)n * n * ncolsC, [&](int j) {
            const double *p = matP + (size_t)ldP * j;
            double *w = W + (size_t)ldW * j;
            for (int i = 0; i < n; ++i) w[i] = 0.0;
            for (int c = 0; c < n; ++c) {
                const double *a = matA + (size_t)ldA * c;
                const double pc = p[c];
                if (ntluA == 'L') {
                    // column c holds A(c..n-1, c); A(c, r) = A(r, c) for r > c
                    w[c] += a[c] * pc;
                    for (int r = c + 1; r < n; ++r) w[r] += a[r] * pc;
                    w[c] += dot(a + c + 1, p + c + 1, n - c - 1);
                } else {
                    for (int r = 0; r < c; ++r) w[r] += a[r] * pc;
                    w[c] += a[c] * pc + dot(a, p, c);
                }
            }
        });
    } else if (ntluA == 'T') {
        gemm_tn(nrowsA, ncolsC, ncolsA, 1.0, matA, ldA, matP, ldP, 0.0, W, ldW);
    } else {
        gemm_nn(nrowsA, ncolsC, ncolsA, matA, ldA, matP, ldP, W, ldW);
    }
    gcge_host_DenseMatQtAP(ntluA, nsdC, nrowsA, nrowsA, nrowsC, ncolsC, alpha, matQ,
                           ldQ, nullptr, ldA, W, ldW, beta, matC, ldC, dbl_ws);
}

}  // extern "C"

namespace {

void hd_MultiVecCreateByVec(void ***des, int num_vec, void *src, OPS *)
{
    mv *s = as_mv(src);
    mv *d = new mv;
    d->nrows = s->nrows;
    d->ncols = num_vec;
    d->ldd = s->ldd;
    d->data = static_cast<double *>(calloc((size_t)d->ldd * num_vec + 1, sizeof(double)));
    *des = reinterpret_cast<void **>(d);
}

void hd_MultiVecCreateByMat(void ***des, int num_vec, void *src_mat, OPS *)
{
    mv *s = as_mv(src_mat);
    mv *d = new mv;
    d->nrows = s->ncols;
    d->ncols = num_vec;
    d->ldd = d->nrows;
    d->data = static_cast<double *>(calloc((size_t)d->ldd * num_vec + 1, sizeof(double)));
    *des = reinterpret_cast<void **>(d);
}

void hd_MultiVecCreateByMultiVec(void ***des, int num_vec, void **src, OPS *ops)
{
    hd_MultiVecCreateByVec(des, num_vec, reinterpret_cast<void *>(src), ops);
}

void hd_MultiVecDestroy(void ***des, int, OPS *)
{
    mv *d = as_mv(*des);
    if (d) {
        free(d->data);
        delete d;
    }
    *des = nullptr;
}

void hd_GetVecFromMultiVec(void **multi_vec, int c, void **vec, OPS *)
{
    mv *m = as_mv(multi_vec);
    mv *v = new mv;
    v->nrows = m->nrows;
    v->ncols = 1;
    v->ldd = m->ldd;
    v->data = col(m, c);
    *vec = v;
}

void hd_RestoreVecForMultiVec(void **, int, void **vec, OPS *)
{
    delete as_mv(*vec);
    *vec = nullptr;
}

void hd_MultiVecView(void **x, int start, int end, OPS *ops)
{
    mv *m = as_mv(x);
    for (int r = 0; r < m->nrows; ++r) {
        for (int c = start; c < end; ++c) ops->Printf("%6.4e\t", col(m, c)[r]);
        ops->Printf("\n");
    }
}

void hd_MultiVecLocalInnerProd(char nsdIP, void **x, void **y, int, int *start, int *end,
                               double *ip, int ldIP, OPS *)
{
    mv *X = as_mv(x), *Y = as_mv(y);
    const int nr = end[0] - start[0], nc = end[1] - start[1];
    if (nr > 0 && nc > 0)
        gcge_host_DenseMatQtAP('S', nsdIP, X->nrows, Y->nrows, nr, nc, 1.0, col(X, start[0]),
                               X->ldd, nullptr, 0, col(Y, start[1]), Y->ldd, 0.0, ip, ldIP,
                               nullptr);
}

void hd_MultiVecSetRandomValue(void **x, int start, int end, OPS *)
{
    mv *m = as_mv(x);
    for (int c = start; c < end; ++c) {
        double *d = col(m, c);
        for (int r = 0; r < m->nrows; ++r) d[r] = (double)rand() / ((double)RAND_MAX + 1.0);
    }
}

void hd_MultiVecAxpby(double alpha, void **x, double beta, void **y, int *start, int *end,
                      OPS *)
{
    assert(end[0] - start[0] == end[1] - start[1]);
    mv *X = as_mv(x), *Y = as_mv(y);
    const int nc = end[1] - start[1];
    if (nc == 0 || Y->nrows == 0) return;
    for (int c = 0; c < nc; ++c) {
        double *d = col(Y, start[1] + c);
        if (beta == 0.0)
            memset(d, 0, sizeof(double) * Y->nrows);
        else if (beta != 1.0)
            for (int r = 0; r < Y->nrows; ++r) d[r] *= beta;
        if (X != nullptr) {
            const double *s = col(X, start[0] + c);
            for (int r = 0; r < Y->nrows; ++r) d[r] += alpha * s[r];
        }
    }
}

void hd_MultiVecLinearComb(void **x, void **y, int, int *start, int *end, double *coef,
                           int ldc, double *beta, int incb, OPS *)
{
    mv *X = as_mv(x), *Y = as_mv(y);
    const int nr = end[0] - start[0], nc = end[1] - start[1];
    if (nr == 0 || nc == 0 || Y->nrows == 0) return;
    const int n = Y->nrows;
    // y_k <- beta_k y_k (beta NULL: overwritten below)
    if (beta != nullptr) {
        for (int k = 0; k < nc; ++k) {
            const double b = (incb == 0) ? beta[0] : beta[(size_t)incb * k];
            if (b == 1.0) continue;
            double *d = col(Y, start[1] + k);
            for (int r = 0; r < n; ++r) d[r] *= b;
        }
    }
    if (X != nullptr && coef != nullptr) {
        std::vector<double> tmp((size_t)n);
        for (int k = 0; k < nc; ++k) {
            for (int r = 0; r < n; ++r) tmp[r] = 0.0;
            for (int i = 0; i < nr; ++i) {
                const double c = coef[(size_t)ldc * k + i];
                const double *s = col(X, start[0] + i);
                for (int r = 0; r < n; ++r) tmp[r] += s[r] * c;
            }
            double *d = col(Y, start[1] + k);
            if (beta == nullptr)
                for (int r = 0; r < n; ++r) d[r] = tmp[r];
            else
                for (int r = 0; r < n; ++r) d[r] += tmp[r];
        }
    }
}

void hd_MatDotMultiVec(void *mat, void **x, void **y, int *start, int *end, OPS *)
{
    assert(end[0] - start[0] == end[1] - start[1]);
    mv *X = as_mv(x), *Y = as_mv(y);
    const int nc = end[1] - start[1];
    if (nc == 0) return;
    if (mat == nullptr) {
        for (int c = 0; c < nc; ++c)
            memcpy(col(Y, start[1] + c), col(X, start[0] + c), sizeof(double) * Y->nrows);
        return;
    }
    mv *A = as_mv(mat);
    gemm_nn(Y->nrows, nc, X->nrows, A->data, A->ldd, col(X, start[0]), X->ldd,
            col(Y, start[1]), Y->ldd);
}

void hd_MatTransDotMultiVec(void *mat, void **x, void **y, int *start, int *end, OPS *)
{
    assert(end[0] - start[0] == end[1] - start[1]);
    mv *X = as_mv(x), *Y = as_mv(y);
    const int nc = end[1] - start[1];
    if (nc == 0) return;
    mv *A = as_mv(mat);
    gemm_tn(Y->nrows, nc, X->nrows, 1.0, A->data, A->ldd, col(X, start[0]), X->ldd, 0.0,
            col(Y, start[1]), Y->ldd);
}

// Q^T op(A) P with the mv_ws side effect (src/ops_multi_vec.c:351-411).
void hd_MultiVecQtAP(char ntsA, char ntsdQAP, void **mvQ, void *matA, void **mvP,
                     int is_vec, int *startQP, int *endQP, double *qAp, int ldQAP,
                     void **mv_ws, OPS *ops)
{
    const int nr = endQP[0] - startQP[0], nc = endQP[1] - startQP[1];
    if (nr <= 0 || nc <= 0) return;
    int s[2], e[2];
    if (matA == nullptr) {
        if (ntsdQAP == 'T') {
            s[0] = startQP[1]; e[0] = endQP[1];
            s[1] = startQP[0]; e[1] = endQP[0];
            ops->MultiVecInnerProd('N', mvP, mvQ, is_vec, s, e, qAp, ldQAP, ops);
        } else {
            ops->MultiVecInnerProd(ntsdQAP, mvQ, mvP, is_vec, startQP, endQP, qAp, ldQAP, ops);
        }
        return;
    }
    s[0] = startQP[1]; e[0] = endQP[1];
    s[1] = 0; e[1] = nc;
    if (ntsA == 'T')
        ops->MatTransDotMultiVec(matA, mvP, mv_ws, s, e, ops);
    else
        ops->MatDotMultiVec(matA, mvP, mv_ws, s, e, ops);
    if (ntsdQAP == 'T') {
        s[0] = 0; e[0] = nc;
        s[1] = startQP[0]; e[1] = endQP[0];
        ops->MultiVecInnerProd('N', mv_ws, mvQ, is_vec, s, e, qAp, ldQAP, ops);
    } else {
        s[0] = startQP[0]; e[0] = endQP[0];
        s[1] = 0; e[1] = nc;
        ops->MultiVecInnerProd(ntsdQAP, mvQ, mv_ws, is_vec, s, e, qAp, ldQAP, ops);
    }
}

void hd_MatView(void *mat, OPS *ops)
{
    mv *A = as_mv(mat);
    for (int r = 0; r < A->nrows; ++r) {
        for (int c = 0; c < A->ncols; ++c) ops->Printf("%6.4e\t", col(A, c)[r]);
        ops->Printf("\n");
    }
}

// single-vector slots: one-column multivectors
void hd_VecCreateByMat(void **v, void *m, OPS *ops)
{
    hd_MultiVecCreateByMat(reinterpret_cast<void ***>(v), 1, m, ops);
}
void hd_VecCreateByVec(void **v, void *s, OPS *ops)
{
    hd_MultiVecCreateByVec(reinterpret_cast<void ***>(v), 1, s, ops);
}
void hd_VecDestroy(void **v, OPS *ops) { hd_MultiVecDestroy(reinterpret_cast<void ***>(v), 1, ops); }
void hd_VecView(void *x, OPS *ops) { hd_MultiVecView(reinterpret_cast<void **>(x), 0, 1, ops); }
void hd_VecInnerProd(void *x, void *y, double *ip, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    hd_MultiVecLocalInnerProd('S', reinterpret_cast<void **>(x), reinterpret_cast<void **>(y),
                              0, s, e, ip, 1, ops);
}
void hd_VecSetRandomValue(void *x, OPS *ops)
{
    hd_MultiVecSetRandomValue(reinterpret_cast<void **>(x), 0, 1, ops);
}
void hd_VecAxpby(double a, void *x, double b, void *y, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    hd_MultiVecAxpby(a, reinterpret_cast<void **>(x), b, reinterpret_cast<void **>(y), s, e,
                     ops);
}
void hd_MatDotVec(void *m, void *x, void *y, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    hd_MatDotMultiVec(m, reinterpret_cast<void **>(x), reinterpret_cast<void **>(y), s, e, ops);
}
void hd_MatTransDotVec(void *m, void *x, void *y, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    hd_MatTransDotMultiVec(m, reinterpret_cast<void **>(x), reinterpret_cast<void **>(y), s, e,
                           ops);
}

void hd_Printf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
}

double hd_GetWtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int hd_GetOption(const char *name, char type, void *value, int argc, char *argv[], OPS *ops)
{
    // "-name value" lookup with the conventions of src/ops_multi_vec.c:58-95
    for (int i = 0; i < argc; ++i) {
        if (strcmp(argv[i], name) != 0) continue;
        if (i + 1 >= argc) return 0;
        const char *v = argv[i + 1];
        if (type == 'i')
            *static_cast<int *>(value) = atoi(v);
        else if (type == 'f')
            *static_cast<double *>(value) = atof(v);
        else if (type == 's')
            strcpy(static_cast<char *>(value), v);
        (void)ops;
        return 1;
    }
    return 0;
}

}  // namespace

static FILE *g_print_target = nullptr;  // stdout unless redirected

extern "C" {

void gcge_host_Printf(const char *fmt, ...)
{
    FILE *f = g_print_target ? g_print_target : stdout;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
    fflush(f);
}

void gcge_host_set_print_target(int to_stderr) { g_print_target = to_stderr ? stderr : nullptr; }

double gcge_host_GetWtime(void) { return hd_GetWtime(); }

int gcge_host_GetOptionFromCommandLine(const char *name, char type, void *value, int argc,
                                       char *argv[], OPS *ops)
{
    return hd_GetOption(name, type, value, argc, argv, ops);
}

// Fills `ops` with the host dense table (the role of OPS_LAPACK_Set,
// app/app_lapack.c:970-1014).
void OPS_HOSTDENSE_Set(OPS *ops)
{
    ops->Printf = gcge_host_Printf;
    ops->GetWtime = gcge_host_GetWtime;
    ops->GetOptionFromCommandLine = gcge_host_GetOptionFromCommandLine;
    ops->MatView = hd_MatView;
    ops->VecCreateByMat = hd_VecCreateByMat;
    ops->VecCreateByVec = hd_VecCreateByVec;
    ops->VecDestroy = hd_VecDestroy;
    ops->VecView = hd_VecView;
    ops->VecInnerProd = hd_VecInnerProd;
    ops->VecLocalInnerProd = hd_VecInnerProd;
    ops->VecSetRandomValue = hd_VecSetRandomValue;
    ops->VecAxpby = hd_VecAxpby;
    ops->MatDotVec = hd_MatDotVec;
    ops->MatTransDotVec = hd_MatTransDotVec;
    ops->MultiVecCreateByMat = hd_MultiVecCreateByMat;
    ops->MultiVecCreateByVec = hd_MultiVecCreateByVec;
    ops->MultiVecCreateByMultiVec = hd_MultiVecCreateByMultiVec;
    ops->MultiVecDestroy = hd_MultiVecDestroy;
    ops->GetVecFromMultiVec = hd_GetVecFromMultiVec;
    ops->RestoreVecForMultiVec = hd_RestoreVecForMultiVec;
    ops->MultiVecView = hd_MultiVecView;
    ops->MultiVecLocalInnerProd = hd_MultiVecLocalInnerProd;
    ops->MultiVecInnerProd = hd_MultiVecLocalInnerProd;
    ops->MultiVecSetRandomValue = hd_MultiVecSetRandomValue;
    ops->MultiVecAxpby = hd_MultiVecAxpby;
    ops->MultiVecLinearComb = hd_MultiVecLinearComb;
    ops->MatDotMultiVec = hd_MatDotMultiVec;
    ops->MatTransDotMultiVec = hd_MatTransDotMultiVec;
    ops->MultiVecQtAP = hd_MultiVecQtAP;
    ops->lapack_ops = nullptr;
    ops->DenseMatQtAP = gcge_host_DenseMatQtAP;
    ops->DenseMatOrth = gcge_host_DenseMatOrth;
}

}  // extern "C"

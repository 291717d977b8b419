// Dense symmetric eigensolvers for the Rayleigh-Ritz step.  Both return
// all eigenvalues ascending in w and orthonormal eigenvectors in Z
// (column-major, ldz); A (lda) is read from its upper triangle and may be
// overwritten.  Return 0 on success.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gcge {
int eigh_host(int n, double *A, int lda, double *w, double *Z, int ldz);
int eigh_device(int n, double *A, int lda, double *w, double *Z, int ldz);
// C (m x n) = A (m x k) B (k x n) + beta C, column-major, on the calling
// thread's stream, by the library GEMM of this thread's context.
void gemm_nn(int64_t m, int n, int k, const double *A, int64_t lda, const double *B, int ldb,
             double beta, double *C, int64_t ldc);
// The same for small device matrices with op(A) = A^T when transa is true
// and any alpha: C (m x n) = alpha op(A) B + beta C.
void gemm_small(bool transa, int m, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc);
// C (n x m, ldc) = A^T, A: m x n (lda), on the device.
void transpose(int m, int n, const double *A, int lda, double *C, int ldc);
}  // namespace gcge

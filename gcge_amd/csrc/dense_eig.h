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
}  // namespace gcge

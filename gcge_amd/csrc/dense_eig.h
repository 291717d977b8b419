// Dense symmetric eigensolvers for the Rayleigh-Ritz step.  Both return
// all eigenvalues ascending in w and orthonormal eigenvectors in Z
// (column-major, ldz); A (lda) is read from its upper triangle and may be
// overwritten.  Return 0 on success.
#pragma once

namespace gcge {
int eigh_host(int n, double *A, int lda, double *w, double *Z, int ldz);
int eigh_device(int n, double *A, int lda, double *w, double *Z, int ldz);
}  // namespace gcge

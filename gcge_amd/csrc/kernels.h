// Launchers of the gfx950 kernels.  All take device pointers and run on
// the given stream; none synchronizes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gcge {

// Sparse matrix in SELL-64: rows grouped in slices of 64 (one wavefront),
// slice s has width w_s = its longest row; entry j of lane l of slice s is
// at slice_ptr[s] + j*64 + l.  Padding entries have value 0 and the row's
// own index as column.
struct SellMat {
    int64_t nrows = 0;
    int64_t nslices = 0;
    int64_t nnz = 0;      // stored, including padding
    int64_t nnz_true = 0; // of the CCS input
    int max_width = 0;
    // slices per grid plane: the typical reach max(col - row) of a row, in
    // whole slices (0: under two slices, no plane structure to exploit)
    int64_t plane = 0;
    // in-plane reach: the typical largest |col - row| below half a plane
    // (rows; 0 when unknown)
    int64_t halo = 0;
    // Constant-coefficient 7-point operator on an n^3 grid (diagonal
    // stencil_d, every in-grid neighbour stencil_o, point x + n y + n^2 z):
    // stencil_n = n when the matrix is one and products may run matrix-free
    // (stencil.hip), 0 otherwise.
    int stencil_n = 0;
    double stencil_d = 0.0, stencil_o = 0.0;
    uint64_t structure_hash = 0;   // of the slice widths, to match A and B
    int64_t *slice_ptr = nullptr;  // nslices + 1
    int *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
};

// Sets A.plane and A.halo from the offsets of a sample of rows
// (stream-ordered, waits for the stream).
void detect_plane(SellMat &A, hipStream_t s);
// Largest grid side of the matrix-free stencil kernels (a wavefront holds
// one x-line, two points per lane).
constexpr int kStencilMaxN = 128;
// Y[:, c] = A X[:, c] by the matrix-free stencil kernel when A is a
// constant-coefficient 7-point operator it handles (false otherwise; the
// same bits as the SELL kernels).
bool spmm_stencil(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
                  int ncols, hipStream_t s);
// R[:, c] = A X[:, c] - lam[c] B X[:, c] (lam on the device) when A and B
// are constant-coefficient 7-point operators on the same grid that the
// stencil kernels handle (false otherwise): one pass over X, and the bits
// that spmm by A and by B, a scaling of the second by lam (scale_cols) and
// their difference (axpby) give.
bool stencil_residual(const SellMat &A, const SellMat &B, const double *X, int64_t ldx,
                      const double *lam, double *R, int64_t ldr, int ncols, hipStream_t s);
// The FP32 BlockPCG product w = A p over the listed 8-column tiles of its
// interleaved blocks, matrix-free for a stencil operator (unshifted), with
// partial p.w sums: stencil_pw_parts(A, nt_bound) per column for a launch
// over at most nt_bound tiles (0 when A is not a stencil operator), at
// partial[(tile 8 + t) parts + part].  The values are the FP32 copies of
// the stencil's, as the matrix's FP32 copy holds.
int stencil_pw_parts(const SellMat &A, int nt_bound);
void stencil_pw(const SellMat &A, const int *tiles, const int *ntact_p, int nt_bound,
                const float *P, float *W, float *partial, hipStream_t s);
// Y[:, c] = A X[:, c], c < ncols.
void spmm(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
          int ncols, hipStream_t s);
// Y = alpha X + beta Y over an nrows x ncols block; X null scales only;
// beta == 0 overwrites Y without reading it.
void axpby(int64_t nrows, int ncols, double alpha, const double *X, int64_t ldx,
           double beta, double *Y, int64_t ldy, hipStream_t s);
// Y[:, c] *= beta[c] (device array) -- or beta_scalar when beta == null.
void scale_cols(int64_t nrows, int ncols, const double *beta, double beta_scalar,
                double *Y, int64_t ldy, hipStream_t s);
// Y[:, c] = X[:, c] (nrows x ncols).
void copy_cols(int64_t nrows, int ncols, const double *X, int64_t ldx, double *Y,
               int64_t ldy, hipStream_t s);
// out[c] = X[:, c] . Y[:, c]; partial is scratch of dots_scratch(ncols) doubles.
int64_t dots_scratch(int ncols);
void col_dots(int64_t nrows, int ncols, const double *X, int64_t ldx, const double *Y,
              int64_t ldy, double *out, double *partial, hipStream_t s);
// G (m x n, ldg) = X^T Y with X: nrows x m, Y: nrows x n.  lower_only
// computes only entries with row >= col (the 'S' case).  The plan fixes
// the kernel, its row chunking and the doubles of partial-sum scratch it
// needs; it is computed once per call and used both to size the scratch
// and to launch, so a change of the tuner table by another thread between
// the two cannot make the launch need more scratch than was sized.
struct GramPlan {
    bool thin = false;    // streaming kernel (a side of at most 8 columns)
    int choice = -1;      // matrix-core kernel: row-chunking candidate, or -1:
    int64_t max_chunks = 0;  // at most this many row chunks (choice -1)
    bool tune = false;    // time every chunking first, on a private buffer
    int64_t scratch = 0;  // doubles of partial sums (0 or less: nothing fits)
};
GramPlan gram_plan(int64_t nrows, int m, int n, bool lower_only);
// A plan whose partial sums fit in max_scratch doubles and that never
// tunes (the table's choice when it fits, else as many row chunks as fit);
// scratch <= 0 when not even one chunk fits.
GramPlan gram_plan_bounded(int64_t nrows, int m, int n, bool lower_only, int64_t max_scratch);
void gram(const GramPlan &p, int64_t nrows, int m, int n, const double *X, int64_t ldx,
          const double *Y, int64_t ldy, double *G, int ldg, bool lower_only, double *partial,
          hipStream_t s);
// Y (nrows x n) = X (nrows x m) C (m x n) + Y diag(beta), with C given
// transposed and k-major: Ct[k * ldct + j] = C[k][j] (device).
// beta_mode: 0 overwrite, 1 scalar beta_scalar, 2 per-column beta (device).
void lincomb(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Ct,
             int ldct, int beta_mode, double beta_scalar, const double *beta, double *Y,
             int64_t ldy, hipStream_t s);
// The matrix-core implementations behind gram() and lincomb().  The Gram
// takes a row-chunking candidate (gram_mfma_choice), or choice -1 with a
// bound on the number of row chunks, and its partial-sum size.
constexpr int kGramMfmaCands = 10;
int gram_mfma_choice(int64_t nrows, int m, int n, bool lower_only, bool *tune);
int64_t gram_mfma_scratch(int64_t nrows, int m, int n, int choice, int64_t max_chunks = 0);
// doubles of partial sums per row chunk
int64_t gram_mfma_chunk_doubles(int m, int n);
void gram_mfma(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Y,
               int64_t ldy, double *G, int ldg, bool lower_only, int choice, int64_t max_chunks,
               bool tune, double *partial, hipStream_t s);
void lincomb_mfma(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Ct,
                  int ldct, int beta_mode, double beta_scalar, const double *beta, double *Y,
                  int64_t ldy, hipStream_t s);
// Block Cholesky QR coefficients from the lower triangle of G = X^T B X
// (n x n, ldg): C = D^{-1} R^{-1} with D = diag(G)^{1/2} and
// D^{-1} G D^{-1} = R^T R, column-major (C[i + n k]) when colmajor, else
// transposed and k-major (C[i n + k]) as lincomb takes it.  flag[0] = 1
// and C = I when a column's B-norm is below zero_tol or a pivot is not
// above rel_tol times the largest scaled diagonal entry, else 0.  work
// holds cholqr_coef_work(n) doubles.  False (nothing launched) unless
// 1 <= n <= 1024.
int64_t cholqr_coef_work(int n);
bool cholqr_coef(int n, const double *G, int ldg, double zero_tol, double rel_tol,
                 double *work, double *coef, bool colmajor, double *flag, hipStream_t s);
// U[0,1) from SplitMix64(seed, element index) into an nrows x ncols block
// whose first element has global index offset.
void random_uniform(int64_t nrows, int ncols, double *Y, int64_t ldy, uint64_t seed,
                    uint64_t offset, hipStream_t s);
// FP64 <-> FP32 conversions of an nrows x ncols block.
void to_f32(int64_t nrows, int ncols, const double *X, int64_t ldx, float *Y,
            int64_t ldy, hipStream_t s);
void to_f64(int64_t nrows, int ncols, const float *X, int64_t ldx, double *Y,
            int64_t ldy, hipStream_t s);

}  // namespace gcge

// gfx950 kernels of the GCG eigensolver backend.
//
// Layout in HBM: multivectors are column-major (element (r, c) at
// base + r + c*ld, ld = nrows), the layout of LAPACKVEC
// (app/app_lapack.h:17-20).  Sparse matrices are SELL-64 (see kernels.h).
//
// Every kernel is written for 64-wide wavefronts and 256-thread
// workgroups.  Workgroup ids are remapped so that each of the 8 XCDs
// (hardware places workgroup i on XCD i % 8) works on a contiguous range
// of rows: row neighbours of a stencil, and the column tiles of one row
// tile, then share that XCD's L2.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.h"

namespace gcge {
namespace {

constexpr int kWG = 256;
constexpr int kXCD = 8;

__device__ __forceinline__ int64_t xcd_swizzle(int64_t bid, int64_t nblk)
{
    const int64_t q = nblk / kXCD, r = nblk % kXCD;
    const int64_t x = bid % kXCD, k = bid / kXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

inline unsigned grid1d(int64_t n, int per_thread = 1)
{
    int64_t b = (n + (int64_t)kWG * per_thread - 1) / ((int64_t)kWG * per_thread);
    return (unsigned)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------------
// SpMM, SELL-64 row gather.  One thread per row keeps the row's entries
// in registers (W = compile-time bound on the slice width) and sweeps all
// columns of the block in tiles of T, so the matrix is read once per call.
// ---------------------------------------------------------------------
template <typename V, int W, int T>
__global__ __launch_bounds__(kWG) void k_spmm_sell(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ col,
    const V *__restrict__ val, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y,
    int64_t ldy, int ncols)
{
    const int64_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t row = blk * kWG + threadIdx.x;
    if (row >= nrows) return;
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) >> 6);
    int cj[W];
    V vj[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {
            cj[j] = col[base + (int64_t)j * 64 + lane];
            vj[j] = val[base + (int64_t)j * 64 + lane];
        } else {
            cj[j] = (int)row;
            vj[j] = V(0);
        }
    }
    int c0 = 0;
    for (; c0 + T <= ncols; c0 += T) {
        V acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = V(0);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (j < w) {
                const V *xp = X + cj[j] + (int64_t)c0 * ldx;
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t] += vj[j] * xp[(int64_t)t * ldx];
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t) Y[row + (int64_t)(c0 + t) * ldy] = acc[t];
    }
    for (; c0 < ncols; ++c0) {
        V acc = V(0);
#pragma unroll
        for (int j = 0; j < W; ++j)
            if (j < w) acc += vj[j] * X[cj[j] + (int64_t)c0 * ldx];
        Y[row + (int64_t)c0 * ldy] = acc;
    }
}

// Rows of any width: entries re-read from memory for each column.
template <typename V>
__global__ __launch_bounds__(kWG) void k_spmm_sell_wide(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ col,
    const V *__restrict__ val, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y,
    int64_t ldy, int ncols)
{
    const int64_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t row = blk * kWG + threadIdx.x;
    if (row >= nrows) return;
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) >> 6);
    for (int c = 0; c < ncols; ++c) {
        V acc = V(0);
        const V *xc = X + (int64_t)c * ldx;
        for (int j = 0; j < w; ++j) {
            const int64_t e = base + (int64_t)j * 64 + lane;
            acc += val[e] * xc[col[e]];
        }
        Y[row + (int64_t)c * ldy] = acc;
    }
}

// y = A x + sigma B x, A and B with the same slice widths.
template <typename V, int W, int T>
__global__ __launch_bounds__(kWG) void k_spmm_shift_sell(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const V *__restrict__ valA, const int *__restrict__ colB, const V *__restrict__ valB,
    V sigma, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y, int64_t ldy,
    int ncols)
{
    const int64_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t row = blk * kWG + threadIdx.x;
    if (row >= nrows) return;
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) >> 6);
    const bool same = (colA == colB);
    int ca[W], cb[W];
    V va[W], vb[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {
            const int64_t e = base + (int64_t)j * 64 + lane;
            ca[j] = colA[e];
            va[j] = valA[e];
            vb[j] = sigma * valB[e];
            cb[j] = same ? ca[j] : colB[e];
        } else {
            ca[j] = cb[j] = (int)row;
            va[j] = vb[j] = V(0);
        }
    }
    for (int c0 = 0; c0 < ncols; c0 += T) {
        V acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = V(0);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (j < w) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    if (c0 + t < ncols) {
                        const V *xc = X + (int64_t)(c0 + t) * ldx;
                        if (same)
                            acc[t] += (va[j] + vb[j]) * xc[ca[j]];
                        else
                            acc[t] += va[j] * xc[ca[j]] + vb[j] * xc[cb[j]];
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t)
            if (c0 + t < ncols) Y[row + (int64_t)(c0 + t) * ldy] = acc[t];
    }
}

template <typename V, int T>
void launch_spmm(const SellMat &A, const V *val, const V *X, int64_t ldx, V *Y,
                 int64_t ldy, int ncols, hipStream_t s)
{
    if (ncols <= 0 || A.nrows == 0) return;
    const unsigned g = grid1d(A.nrows);
    if (A.max_width <= 8)
        k_spmm_sell<V, 8, T><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X, ldx, Y,
                                               ldy, ncols);
    else if (A.max_width <= 16)
        k_spmm_sell<V, 16, T><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X, ldx, Y,
                                                ldy, ncols);
    else if (A.max_width <= 32)
        k_spmm_sell<V, 32, T><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X, ldx, Y,
                                                ldy, ncols);
    else
        k_spmm_sell_wide<V><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X, ldx, Y,
                                              ldy, ncols);
}

// ---------------------------------------------------------------------
// BLAS-1 over column blocks
// ---------------------------------------------------------------------
constexpr int kRowsPerThread = 4;

__global__ __launch_bounds__(kWG) void k_axpby(int64_t nrows, double alpha,
                                               const double *__restrict__ X, int64_t ldx,
                                               double beta, double *__restrict__ Y,
                                               int64_t ldy)
{
    const int c = blockIdx.y;
    const double *x = X ? X + (int64_t)c * ldx : nullptr;
    double *y = Y + (int64_t)c * ldy;
    const int64_t r0 = (int64_t)blockIdx.x * kWG * kRowsPerThread + threadIdx.x;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
        const int64_t r = r0 + (int64_t)q * kWG;
        if (r < nrows) {
            double v = (beta == 0.0) ? 0.0 : beta * y[r];
            if (x) v += alpha * x[r];
            y[r] = v;
        }
    }
}

__global__ __launch_bounds__(kWG) void k_scale_cols(int64_t nrows,
                                                    const double *__restrict__ beta,
                                                    double bs, double *__restrict__ Y,
                                                    int64_t ldy)
{
    const int c = blockIdx.y;
    const double b = beta ? beta[c] : bs;
    double *y = Y + (int64_t)c * ldy;
    const int64_t r0 = (int64_t)blockIdx.x * kWG * kRowsPerThread + threadIdx.x;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
        const int64_t r = r0 + (int64_t)q * kWG;
        if (r < nrows) y[r] = (b == 0.0) ? 0.0 : b * y[r];
    }
}

__global__ __launch_bounds__(kWG) void k_copy_cols(int64_t nrows,
                                                   const double *__restrict__ X, int64_t ldx,
                                                   double *__restrict__ Y, int64_t ldy)
{
    const int c = blockIdx.y;
    const double *x = X + (int64_t)c * ldx;
    double *y = Y + (int64_t)c * ldy;
    const int64_t r0 = (int64_t)blockIdx.x * kWG * kRowsPerThread + threadIdx.x;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
        const int64_t r = r0 + (int64_t)q * kWG;
        if (r < nrows) y[r] = x[r];
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum of a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kWG / 64];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kWG / 64; ++i) t += red[i];
    }
    return t;
}

constexpr int kDotBlocks = 128;

__global__ __launch_bounds__(kWG) void k_col_dots_partial(
    int64_t nrows, const double *__restrict__ X, int64_t ldx, const double *__restrict__ Y,
    int64_t ldy, double *__restrict__ partial)
{
    const int c = blockIdx.y;
    const double *x = X + (int64_t)c * ldx;
    const double *y = Y + (int64_t)c * ldy;
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * kWG)
        acc += x[r] * y[r];
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[(int64_t)c * gridDim.x + blockIdx.x] = acc;
}

__global__ __launch_bounds__(64) void k_sum_partials(int nparts,
                                                     const double *__restrict__ partial,
                                                     double *__restrict__ out)
{
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) acc += partial[(int64_t)c * nparts + i];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) out[c] = acc;
}

// ---------------------------------------------------------------------
// Gram G = X^T Y: 64 x 64 output tiles, each workgroup owns one tile over
// one chunk of rows and writes a partial tile; a second kernel sums the
// chunks in a fixed order (deterministic, no atomics).
// ---------------------------------------------------------------------
constexpr int kGT = 64;   // output tile
constexpr int kGK = 32;   // rows per LDS stage

struct GramPlan {
    int tm, tn, nchunk;
    int64_t rows_per_chunk;
};

GramPlan gram_plan(int64_t nrows, int m, int n)
{
    GramPlan p;
    p.tm = (m + kGT - 1) / kGT;
    p.tn = (n + kGT - 1) / kGT;
    const int64_t tiles = (int64_t)p.tm * p.tn;
    int64_t want = (2048 + tiles - 1) / tiles;
    int64_t cap = (nrows + 511) / 512;
    int64_t nc = want < cap ? want : cap;
    if (nc < 1) nc = 1;
    if (nc > 1024) nc = 1024;
    int64_t rpc = (nrows + nc - 1) / nc;
    rpc = (rpc + kGK - 1) / kGK * kGK;
    p.rows_per_chunk = rpc;
    p.nchunk = (int)((nrows + rpc - 1) / rpc);
    if (p.nchunk < 1) p.nchunk = 1;
    return p;
}

__global__ __launch_bounds__(kWG) void k_gram_partial(
    int64_t nrows, int m, int n, const double *__restrict__ X, int64_t ldx,
    const double *__restrict__ Y, int64_t ldy, int tm, int tn, int64_t rows_per_chunk,
    int lower_only, double *__restrict__ partial)
{
    __shared__ double sx[kGK][kGT + 1];
    __shared__ double sy[kGK][kGT + 1];
    const int64_t nblk = (int64_t)gridDim.x;
    const int64_t lb = xcd_swizzle(blockIdx.x, nblk);
    const int tiles = tm * tn;
    const int chunk = (int)(lb / tiles);
    const int t = (int)(lb % tiles);
    const int bi = t / tn, bj = t % tn;
    double *out = partial + ((int64_t)chunk * tiles + t) * (kGT * kGT);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (lower_only && bi < bj) {
        // tile strictly above the diagonal: contributes nothing
        for (int e = threadIdx.x; e < kGT * kGT; e += kWG) out[e] = 0.0;
        return;
    }
    const int i0 = bi * kGT, j0 = bj * kGT;
    const int64_t rbeg = (int64_t)chunk * rows_per_chunk;
    int64_t rend = rbeg + rows_per_chunk;
    if (rend > nrows) rend = nrows;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int64_t r0 = rbeg; r0 < rend; r0 += kGK) {
        // stage X[r0:r0+32, i0:i0+64] and Y[r0:r0+32, j0:j0+64]
#pragma unroll
        for (int q = 0; q < (kGK * kGT) / kWG; ++q) {
            const int e = threadIdx.x + q * kWG;
            const int rr = e % kGK, cc = e / kGK;
            const int64_t r = r0 + rr;
            const bool rok = r < rend;
            sx[rr][cc] = (rok && i0 + cc < m) ? X[r + (int64_t)(i0 + cc) * ldx] : 0.0;
            sy[rr][cc] = (rok && j0 + cc < n) ? Y[r + (int64_t)(j0 + cc) * ldy] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kGK; ++k) {
            double xa[4], yb[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) xa[a] = sx[k][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) yb[b] = sy[k][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += xa[a] * yb[b];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) out[(ty + 16 * a) * kGT + tx + 16 * b] = acc[a][b];
}

__global__ __launch_bounds__(kWG) void k_gram_reduce(int m, int n, int tm, int tn,
                                                    int nchunk,
                                                    const double *__restrict__ partial,
                                                    double *__restrict__ G, int ldg)
{
    const int64_t idx = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (idx >= (int64_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    const int bi = i / kGT, bj = j / kGT;
    const int t = bi * tn + bj;
    const int64_t off = (int64_t)t * (kGT * kGT) + (i % kGT) * kGT + (j % kGT);
    const int64_t stride = (int64_t)tm * tn * kGT * kGT;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[off + (int64_t)c * stride];
    G[i + (int64_t)j * ldg] = s;
}

// ---------------------------------------------------------------------
// misc
// ---------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kWG) void k_random(int64_t nrows, double *__restrict__ Y,
                                                int64_t ldy, uint64_t seed, uint64_t offset)
{
    const int c = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r >= nrows) return;
    const uint64_t idx = offset + (uint64_t)c * (uint64_t)nrows + (uint64_t)r;
    const uint64_t z = splitmix64(seed ^ splitmix64(idx));
    Y[r + (int64_t)c * ldy] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kWG) void k_to_f32(int64_t nrows, const double *__restrict__ X,
                                                int64_t ldx, float *__restrict__ Y,
                                                int64_t ldy)
{
    const int c = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r < nrows) Y[r + (int64_t)c * ldy] = (float)X[r + (int64_t)c * ldx];
}

__global__ __launch_bounds__(kWG) void k_to_f64(int64_t nrows, const float *__restrict__ X,
                                                int64_t ldx, double *__restrict__ Y,
                                                int64_t ldy)
{
    const int c = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r < nrows) Y[r + (int64_t)c * ldy] = (double)X[r + (int64_t)c * ldx];
}

}  // namespace

void spmm(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy, int ncols,
          hipStream_t s)
{
    launch_spmm<double, 8>(A, A.val, X, ldx, Y, ldy, ncols, s);
}

void spmm32(const SellMat &A, const float *X, int64_t ldx, float *Y, int64_t ldy, int ncols,
            hipStream_t s)
{
    launch_spmm<float, 16>(A, A.val32, X, ldx, Y, ldy, ncols, s);
}

void spmm_shift32(const SellMat &A, const SellMat &B, float sigma, const float *X,
                  int64_t ldx, float *Y, int64_t ldy, int ncols, hipStream_t s)
{
    if (ncols <= 0 || A.nrows == 0) return;
    const unsigned g = grid1d(A.nrows);
    if (A.max_width <= 8)
        k_spmm_shift_sell<float, 8, 8><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col,
                                                         A.val32, B.col, B.val32, sigma, X,
                                                         ldx, Y, ldy, ncols);
    else
        k_spmm_shift_sell<float, 32, 4><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col,
                                                          A.val32, B.col, B.val32, sigma, X,
                                                          ldx, Y, ldy, ncols);
}

void axpby(int64_t nrows, int ncols, double alpha, const double *X, int64_t ldx, double beta,
           double *Y, int64_t ldy, hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows, kRowsPerThread), ncols);
    k_axpby<<<g, kWG, 0, s>>>(nrows, alpha, X, ldx, beta, Y, ldy);
}

void scale_cols(int64_t nrows, int ncols, const double *beta, double beta_scalar, double *Y,
                int64_t ldy, hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows, kRowsPerThread), ncols);
    k_scale_cols<<<g, kWG, 0, s>>>(nrows, beta, beta_scalar, Y, ldy);
}

void copy_cols(int64_t nrows, int ncols, const double *X, int64_t ldx, double *Y,
               int64_t ldy, hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows, kRowsPerThread), ncols);
    k_copy_cols<<<g, kWG, 0, s>>>(nrows, X, ldx, Y, ldy);
}

int64_t dots_scratch(int ncols) { return (int64_t)ncols * kDotBlocks; }

void col_dots(int64_t nrows, int ncols, const double *X, int64_t ldx, const double *Y,
              int64_t ldy, double *out, double *partial, hipStream_t s)
{
    if (ncols <= 0) return;
    int nb = (int)((nrows + kWG * 8 - 1) / (kWG * 8));
    if (nb < 1) nb = 1;
    if (nb > kDotBlocks) nb = kDotBlocks;
    dim3 g(nb, ncols);
    k_col_dots_partial<<<g, kWG, 0, s>>>(nrows, X, ldx, Y, ldy, partial);
    k_sum_partials<<<ncols, 64, 0, s>>>(nb, partial, out);
}

// Blocks of at least 16 x 16 go to the matrix cores; thinner ones (the
// one-column products of the column-by-column orthogonalization) to the
// FMA kernel.
static bool gram_use_mfma(int m, int n) { return m >= 16 && n >= 16; }

int64_t gram_scratch(int64_t nrows, int m, int n)
{
    if (gram_use_mfma(m, n)) return gram_mfma_scratch(nrows, m, n);
    GramPlan p = gram_plan(nrows, m, n);
    return (int64_t)p.nchunk * p.tm * p.tn * kGT * kGT;
}

void gram(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Y,
          int64_t ldy, double *G, int ldg, bool lower_only, double *partial, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    if (gram_use_mfma(m, n)) {
        gram_mfma(nrows, m, n, X, ldx, Y, ldy, G, ldg, lower_only, partial, s);
        return;
    }
    GramPlan p = gram_plan(nrows, m, n);
    const int64_t nblk = (int64_t)p.nchunk * p.tm * p.tn;
    k_gram_partial<<<(unsigned)nblk, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.tm, p.tn,
                                                  p.rows_per_chunk, lower_only ? 1 : 0,
                                                  partial);
    k_gram_reduce<<<grid1d((int64_t)m * n), kWG, 0, s>>>(m, n, p.tm, p.tn, p.nchunk, partial,
                                                        G, ldg);
}

void lincomb(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Ct,
             int ldct, int beta_mode, double beta_scalar, const double *beta, double *Y,
             int64_t ldy, hipStream_t s)
{
    lincomb_mfma(nrows, m, n, X, ldx, Ct, ldct, beta_mode, beta_scalar, beta, Y, ldy, s);
}

void random_uniform(int64_t nrows, int ncols, double *Y, int64_t ldy, uint64_t seed,
                    uint64_t offset, hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows), ncols);
    k_random<<<g, kWG, 0, s>>>(nrows, Y, ldy, seed, offset);
}

void to_f32(int64_t nrows, int ncols, const double *X, int64_t ldx, float *Y, int64_t ldy,
            hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows), ncols);
    k_to_f32<<<g, kWG, 0, s>>>(nrows, X, ldx, Y, ldy);
}

void to_f64(int64_t nrows, int ncols, const float *X, int64_t ldx, double *Y, int64_t ldy,
            hipStream_t s)
{
    if (nrows <= 0 || ncols <= 0) return;
    dim3 g(grid1d(nrows), ncols);
    k_to_f64<<<g, kWG, 0, s>>>(nrows, X, ldx, Y, ldy);
}

}  // namespace gcge

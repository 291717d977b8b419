// gfx950 kernels of the GCG eigensolver backend.
//
// Layout in HBM: multivectors are column-major (element (r, c) at
// base + r + c*ld, ld = nrows), the layout of LAPACKVEC
// (app/app_lapack.h:17-20).  Sparse matrices are SELL-64 (see kernels.h).
//
// Every kernel is written for 64-wide wavefronts and 256-thread
// workgroups and takes workgroups in plain order (the hardware places
// workgroup i on XCD i % 8, so row blocks interleave over the XCDs; giving
// each XCD a contiguous row range instead measured ~10 % slower for the
// SpMM gather).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <vector>

#include "app_hip_internal.h"
#include "kernels.h"
#include "sell_rows.h"
#include "tuner.h"

namespace gcge {
namespace {

constexpr int kWG = 256;

inline unsigned grid1d(int64_t n, int per_thread = 1)
{
    int64_t b = (n + (int64_t)kWG * per_thread - 1) / ((int64_t)kWG * per_thread);
    return (unsigned)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------------
// SpMM, SELL-64 row gather.  One thread per row keeps the row's entries
// in registers (W = compile-time bound on the slice width) and sweeps all
// columns of the block in tiles of T, so the matrix is read once per call.
// The tile loop runs in a body compiled for the slice's exact width
// (for_width), so all w x T gathers of a tile are in flight together.
// ---------------------------------------------------------------------
// Row-block orders: plain, one contiguous range per XCD, plane-grouped,
// plane-grouped over XCD-contiguous blocks; kLds: plain order with the X
// window staged in LDS (k_spmm_sell_lds).
enum RowOrder { kPlain = 0, kXcd = 1, kPlanes = 2, kPlanesXcd = 3, kLds = 4 };
constexpr int kMaxHalo = 1024;

template <typename V, int W, int T, int BS, int ORD>
__global__ __launch_bounds__(BS) void k_spmm_sell(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ col,
    const V *__restrict__ val, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y,
    int64_t ldy, int ncols, int64_t plane)
{
    static_assert(BS % 64 == 0, "one wavefront per slice");
    static_assert(ORD < kPlanes || BS == 256, "four planes per workgroup");
    int64_t s;
    if constexpr (ORD == kPlanes) {
        s = plane_slice(plane, blockIdx.x);
    } else if constexpr (ORD == kPlanesXcd) {
        s = plane_slice(plane, row_block<true>());
    } else {
        s = wave_slice<BS>(row_block<ORD == kXcd>());
    }
    const int64_t row = s * 64 + (threadIdx.x & 63);
    if (row >= nrows) return;
    const int lane = (int)(threadIdx.x & 63);
    const int64_t base = slice_ptr[s];
    const int w = __builtin_amdgcn_readfirstlane((int)((slice_ptr[s + 1] - base) >> 6));
    // gathers address a wave-uniform column base plus a 32-bit byte offset
    // per entry (one VGPR each instead of a 64-bit address per load)
    uint32_t oj[W];
    V vj[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {  // wave-uniform
            oj[j] = (uint32_t)col[base + (int64_t)j * 64 + lane] * (uint32_t)sizeof(V);
            vj[j] = val[base + (int64_t)j * 64 + lane];
        }
    }
    // one buffer resource per column: a column holds nrows values
    const uint32_t colbytes = (uint32_t)(nrows * (int64_t)sizeof(V));
    for_width<W>(w, [&](auto kw) {
        constexpr int K = decltype(kw)::value;
        int c0 = 0;
        for (; c0 + T <= ncols; c0 += T) {
            __amdgpu_buffer_rsrc_t rs[T];
#pragma unroll
            for (int t = 0; t < T; ++t) rs[t] = buffer_rsrc(X + (int64_t)(c0 + t) * ldx, colbytes);
            V acc[T];
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = V(0);
#pragma unroll
            for (int j = 0; j < K; ++j) {
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t] += vj[j] * buffer_load<V>(rs[t], oj[j]);
            }
#pragma unroll
            for (int t = 0; t < T; ++t) Y[row + (int64_t)(c0 + t) * ldy] = acc[t];
        }
        for (; c0 < ncols; ++c0) {
            const __amdgpu_buffer_rsrc_t r1 = buffer_rsrc(X + (int64_t)c0 * ldx, colbytes);
            V acc = V(0);
#pragma unroll
            for (int j = 0; j < K; ++j) acc += vj[j] * buffer_load<V>(r1, oj[j]);
            Y[row + (int64_t)c0 * ldy] = acc;
        }
    });
}

// The same product with the X rows of the workgroup's 256 rows and `halo`
// rows either side staged in LDS per tile of T columns (filled by coalesced
// loads, column-major, one spare slot per column).  Entries whose column
// falls outside the window are loaded from memory first, in a lane-masked
// region whose results are not used inside it (lanes with an in-window
// entry issue no request, and a stencil's interior slices skip the region
// as a whole); then every entry takes its value from LDS or from the
// loaded register, in entry order, so the sums are those of k_spmm_sell.
template <typename V, int W, int T>
__global__ __launch_bounds__(256) void k_spmm_sell_lds(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ col,
    const V *__restrict__ val, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y,
    int64_t ldy, int ncols, int halo)
{
    extern __shared__ unsigned char lds_raw[];
    V *win = reinterpret_cast<V *>(lds_raw);
    const int span = 256 + 2 * halo;
    const int ld = span + 1;  // slot span of each column: read by out-of-window entries
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t lo = r0 - halo;
    const int64_t row = r0 + threadIdx.x;
    const bool live = row < nrows;
    const int64_t s = wave_slice<256>(blockIdx.x);
    const int lane = (int)(threadIdx.x & 63);
    int w = 0;
    int64_t base = 0;
    if (s < (nrows + 63) / 64) {  // wave-uniform
        base = slice_ptr[s];
        w = __builtin_amdgcn_readfirstlane((int)((slice_ptr[s + 1] - base) >> 6));
    }
    const uint32_t colbytes = (uint32_t)(nrows * (int64_t)sizeof(V));
    uint32_t og[W];
    int li[W];
    bool nr[W];
    V vj[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {  // wave-uniform; the slice's padding covers every lane
            const int64_t e = base + (int64_t)j * 64 + lane;
            const int c = col[e];
            vj[j] = val[e];
            const int64_t d = (int64_t)c - lo;
            nr[j] = d >= 0 && d < span;
            og[j] = (uint32_t)c * (uint32_t)sizeof(V);
            li[j] = nr[j] ? (int)d : span;
        }
    }
    for (int c0 = 0; c0 < ncols; c0 += T) {
        const int tn = ncols - c0 < T ? ncols - c0 : T;  // uniform
        for (int i = threadIdx.x; i < span * tn; i += 256) {
            const int t = i / span, rr = i - t * span;
            const int64_t g = lo + rr;
            win[t * ld + rr] = (g >= 0 && g < nrows) ? X[g + (int64_t)(c0 + t) * ldx] : V(0);
        }
        __syncthreads();
        __amdgpu_buffer_rsrc_t rs[T];
#pragma unroll
        for (int t = 0; t < T; ++t)
            rs[t] = buffer_rsrc(X + (int64_t)(c0 + (t < tn ? t : 0)) * ldx, colbytes);
        for_width<W>(w, [&](auto kw) {
            constexpr int K = decltype(kw)::value;
            V xg[K > 0 ? K : 1][T];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (!nr[j]) {
#pragma unroll
                    for (int t = 0; t < T; ++t) xg[j][t] = buffer_load<V>(rs[t], og[j]);
                }
            }
            // every LDS read issued before any is consumed; the empty asm
            // statements keep them unconditional, so the choice below is a
            // select rather than a branch around each read
            V xl[K > 0 ? K : 1][T];
#pragma unroll
            for (int j = 0; j < K; ++j)
#pragma unroll
                for (int t = 0; t < T; ++t) xl[j][t] = win[t * ld + li[j]];
#pragma unroll
            for (int j = 0; j < K; ++j)
#pragma unroll
                for (int t = 0; t < T; ++t) asm volatile("" : "+v"(xl[j][t]));
            V acc[T];
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = V(0);
#pragma unroll
            for (int j = 0; j < K; ++j) {
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t] += vj[j] * (nr[j] ? xl[j][t] : xg[j][t]);
            }
            if (live) {
#pragma unroll
                for (int t = 0; t < T; ++t)
                    if (t < tn) Y[row + (int64_t)(c0 + t) * ldy] = acc[t];
            }
        });
        __syncthreads();
    }
}

// Rows of any width: entries re-read from memory for each column.
template <typename V>
__global__ __launch_bounds__(kWG) void k_spmm_sell_wide(
    int64_t nrows, const int64_t *__restrict__ slice_ptr, const int *__restrict__ col,
    const V *__restrict__ val, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y,
    int64_t ldy, int ncols)
{
    // plain block order (remapping blocks to give each XCD a contiguous row
    // range measured ~10 % slower for this gather)
    const int64_t blk = blockIdx.x;
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

// Candidate launch shapes per entry bound W, the first being the default:
// column tiles narrow as rows widen (W entries x T columns of gathers in
// flight per thread), so registers stay near the same budget -- 4 columns
// for rows of up to 8 entries, 2 up to 16, 1 up to 32 -- in 256-thread
// workgroups; the tuner may pick narrower tiles or 64-thread workgroups
// (more workgroups for small matrices).  All give the same bits.
template <typename V>
using SpmmFn = void (*)(const SellMat &, const V *, const V *, int64_t, V *, int64_t, int,
                        hipStream_t);

template <typename V, int W, int T, int BS, int ORD>
void spmm_shape(const SellMat &A, const V *val, const V *X, int64_t ldx, V *Y, int64_t ldy,
                int ncols, hipStream_t s)
{
    if constexpr (ORD == kLds) {
        if (A.halo > 0 && A.halo <= kMaxHalo) {
            const int halo = (int)A.halo;
            const size_t lds = (size_t)(256 + 2 * halo + 1) * T * sizeof(V);
            k_spmm_sell_lds<V, W, T><<<(unsigned)((A.nrows + 255) / 256), 256, lds, s>>>(
                A.nrows, A.slice_ptr, A.col, val, X, ldx, Y, ldy, ncols, halo);
            return;
        }
        k_spmm_sell<V, W, T, 256, kPlain><<<(unsigned)((A.nrows + 255) / 256), 256, 0, s>>>(
            A.nrows, A.slice_ptr, A.col, val, X, ldx, Y, ldy, ncols, 0);
    } else if constexpr (ORD >= kPlanes) {
        if (A.plane > 0) {
            const int64_t groups = (A.nslices + 4 * A.plane - 1) / (4 * A.plane);
            k_spmm_sell<V, W, T, 256, ORD><<<(unsigned)(groups * A.plane), 256, 0, s>>>(
                A.nrows, A.slice_ptr, A.col, val, X, ldx, Y, ldy, ncols, A.plane);
            return;
        }
        // no plane structure: plain order
        k_spmm_sell<V, W, T, 256, kPlain><<<(unsigned)((A.nrows + 255) / 256), 256, 0, s>>>(
            A.nrows, A.slice_ptr, A.col, val, X, ldx, Y, ldy, ncols, 0);
    } else {
        const unsigned g = (unsigned)((A.nrows + BS - 1) / BS);
        k_spmm_sell<V, W, T, BS, ORD><<<g, BS, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X, ldx,
                                                       Y, ldy, ncols, 0);
    }
}

// (shapes added later come last so that table entries made before they
// existed keep their meaning)
template <typename V>
struct SpmmShapes {
    static constexpr SpmmFn<V> w8[15] = {
        spmm_shape<V, 8, 4, 256, kPlain>, spmm_shape<V, 8, 2, 256, kPlain>,
        spmm_shape<V, 8, 1, 256, kPlain>, spmm_shape<V, 8, 4, 64, kPlain>,
        spmm_shape<V, 8, 2, 64, kPlain>,  spmm_shape<V, 8, 1, 64, kPlain>,
        spmm_shape<V, 8, 4, 256, kXcd>,   spmm_shape<V, 8, 2, 256, kXcd>,
        spmm_shape<V, 8, 1, 256, kXcd>,   spmm_shape<V, 8, 4, 256, kPlanes>,
        spmm_shape<V, 8, 1, 256, kPlanes>, spmm_shape<V, 8, 4, 256, kPlanesXcd>,
        spmm_shape<V, 8, 1, 256, kPlanesXcd>, spmm_shape<V, 8, 4, 256, kLds>,
        spmm_shape<V, 8, 1, 256, kLds>};
    static constexpr SpmmFn<V> w16[11] = {
        spmm_shape<V, 16, 2, 256, kPlain>, spmm_shape<V, 16, 1, 256, kPlain>,
        spmm_shape<V, 16, 2, 64, kPlain>,  spmm_shape<V, 16, 1, 64, kPlain>,
        spmm_shape<V, 16, 2, 256, kXcd>,   spmm_shape<V, 16, 1, 256, kXcd>,
        spmm_shape<V, 16, 2, 256, kPlanes>, spmm_shape<V, 16, 1, 256, kPlanes>,
        spmm_shape<V, 16, 1, 256, kPlanesXcd>, spmm_shape<V, 16, 2, 256, kLds>,
        spmm_shape<V, 16, 1, 256, kLds>};
    static constexpr SpmmFn<V> w32[6] = {
        spmm_shape<V, 32, 1, 256, kPlain>, spmm_shape<V, 32, 1, 64, kPlain>,
        spmm_shape<V, 32, 1, 256, kXcd>, spmm_shape<V, 32, 1, 256, kPlanes>,
        spmm_shape<V, 32, 1, 256, kPlanesXcd>, spmm_shape<V, 32, 1, 256, kLds>};
};

int ncols_class(int ncols)
{
    return ncols <= 1 ? 0 : ncols <= 2 ? 1 : ncols <= 4 ? 2 : ncols <= 8 ? 3 : ncols <= 32 ? 4 : 5;
}

template <typename V>
void launch_spmm(const SellMat &A, const V *val, const V *X, int64_t ldx, V *Y,
                 int64_t ldy, int ncols, hipStream_t s)
{
    if (ncols <= 0 || A.nrows == 0) return;
    // the register kernels gather by 32-bit byte offsets within a column
    const bool small = A.nrows * (int64_t)sizeof(V) <= (int64_t)UINT32_MAX - 64;
    if (!small || A.max_width > 32) {
        k_spmm_sell_wide<V><<<grid1d(A.nrows), kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, val, X,
                                                            ldx, Y, ldy, ncols);
        return;
    }
    const int W = A.max_width <= 8 ? 8 : A.max_width <= 16 ? 16 : 32;
    const SpmmFn<V> *cand = W == 8 ? SpmmShapes<V>::w8 : W == 16 ? SpmmShapes<V>::w16
                                                                 : SpmmShapes<V>::w32;
    const int ncand = W == 8 ? 15 : W == 16 ? 11 : 6;
    const TuneKey key{kTuneSpmm, W + (sizeof(V) == 4 ? 1000 : 0), ilog2_floor(A.nrows),
                      ncols_class(ncols)};
    int ch = tune_lookup(key);
    if (ch < 0 || ch >= ncand) {
        if (tune_enabled()) {
            tune_run(key, ncand, 0,
                     [&](int i) { cand[i](A, val, X, ldx, Y, ldy, ncols, s); }, s);
            return;
        }
        ch = 0;
    }
    cand[ch](A, val, X, ldx, Y, ldy, ncols, s);
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
// Thin Gram blocks, G (m x n) = X^T Y with n <= 8 (the one-column products
// of the column-by-column orthogonalization): a GEMV-like stream, one
// thread per row, MT = 64 / NB columns of X per workgroup and all n columns
// of Y in registers; X and Y are each read once per column tile.  Partial
// sums per row chunk are combined by k_gram_thin_reduce in a fixed order.
// ---------------------------------------------------------------------
struct ThinPlan {
    int nb, mt, mtiles, nchunk;
    int64_t rows_per_chunk;
};

ThinPlan thin_plan(int64_t nrows, int m, int n)
{
    ThinPlan p;
    p.nb = n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : 8;
    p.mt = 64 / p.nb;
    p.mtiles = (m + p.mt - 1) / p.mt;
    // a few long-lived workgroups per CU: the register accumulators are
    // reduced once per workgroup, so each should stream many rows (at most
    // 512 in total, so no straggler round)
    int64_t want = 512 / p.mtiles;
    int64_t cap = (nrows + 1023) / 1024;  // at least 1024 rows per chunk
    int64_t nc = want < cap ? want : cap;
    if (nc < 1) nc = 1;
    p.rows_per_chunk = (nrows + nc - 1) / nc;
    p.nchunk = (int)((nrows + p.rows_per_chunk - 1) / p.rows_per_chunk);
    if (p.nchunk < 1) p.nchunk = 1;
    return p;
}

template <int NB>
__global__ __launch_bounds__(kWG) void k_gram_thin(int64_t nrows, int m, int n,
                                                   const double *__restrict__ X, int64_t ldx,
                                                   const double *__restrict__ Y, int64_t ldy,
                                                   int64_t rows_per_chunk,
                                                   double *__restrict__ partial)
{
    constexpr int MT = 64 / NB;
    __shared__ double red[kWG / 64][MT * NB];
    const int chunk = blockIdx.x, mt = blockIdx.y;
    const int i0 = mt * MT;
    const int64_t rbeg = (int64_t)chunk * rows_per_chunk;
    int64_t rend = rbeg + rows_per_chunk;
    if (rend > nrows) rend = nrows;
    double acc[MT][NB];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int q = 0; q < NB; ++q) acc[i][q] = 0.0;
    for (int64_t r = rbeg + threadIdx.x; r < rend; r += kWG) {
        double yv[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) yv[q] = (q < n) ? Y[r + (int64_t)q * ldy] : 0.0;
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            if (i0 + i < m) {
                const double xv = X[r + (int64_t)(i0 + i) * ldx];
#pragma unroll
                for (int q = 0; q < NB; ++q) acc[i][q] += xv * yv[q];
            }
        }
    }
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const double v = wave_sum(acc[i][q]);
            if (ln == 0) red[wv][i * NB + q] = v;
        }
    __syncthreads();
    for (int e = threadIdx.x; e < MT * NB; e += kWG) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) s += red[w][e];
        partial[((int64_t)chunk * gridDim.y + mt) * (MT * NB) + e] = s;
    }
}

// G[i + j*ldg] (or G[j + i*ldg] when transposed) = sum over chunks
__global__ __launch_bounds__(kWG) void k_gram_thin_reduce(int m, int n, int mt, int nb,
                                                         int mtiles, int nchunk,
                                                         const double *__restrict__ partial,
                                                         double *__restrict__ G, int ldg,
                                                         int transposed)
{
    const int64_t off = (int64_t)blockIdx.x * kWG + threadIdx.x;
    const int64_t stride = (int64_t)mtiles * mt * nb;
    if (off >= stride) return;
    const int tile = (int)(off / (mt * nb)), within = (int)(off % (mt * nb));
    const int i = tile * mt + within / nb, j = within % nb;
    if (i >= m || j >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[off + (int64_t)c * stride];
    if (transposed)
        G[j + (int64_t)i * ldg] = s;
    else
        G[i + (int64_t)j * ldg] = s;
}

void gram_thin(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Y,
               int64_t ldy, double *G, int ldg, bool transposed, double *partial, hipStream_t s)
{
    ThinPlan p = thin_plan(nrows, m, n);
    dim3 g(p.nchunk, p.mtiles);
    switch (p.nb) {
    case 1: k_gram_thin<1><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.rows_per_chunk, partial); break;
    case 2: k_gram_thin<2><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.rows_per_chunk, partial); break;
    case 4: k_gram_thin<4><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.rows_per_chunk, partial); break;
    default: k_gram_thin<8><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.rows_per_chunk, partial); break;
    }
    const int64_t total = (int64_t)p.mtiles * p.mt * p.nb;
    k_gram_thin_reduce<<<grid1d(total), kWG, 0, s>>>(m, n, p.mt, p.nb, p.mtiles, p.nchunk,
                                                     partial, G, ldg, transposed ? 1 : 0);
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

namespace {

// off[k][j] = col - row of entry j < 32 of sampled row k (zero past the
// row's width, which the padding also gives)
__global__ __launch_bounds__(kWG) void k_row_offsets(int64_t nrows, int nsample,
                                                     const int64_t *__restrict__ slice_ptr,
                                                     const int *__restrict__ col,
                                                     int64_t *__restrict__ off)
{
    const int k = blockIdx.x * kWG + threadIdx.x;
    if (k >= nsample) return;
    const int64_t r = (int64_t)k * nrows / nsample;
    const int64_t s = r >> 6;
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) >> 6);
    for (int j = 0; j < 32; ++j)
        off[(int64_t)k * 32 + j] =
            j < w ? (int64_t)col[base + (int64_t)j * 64 + (r & 63)] - r : 0;
}

}  // namespace

void detect_plane(SellMat &A, hipStream_t s)
{
    A.plane = 0;
    A.halo = 0;
    if (A.nrows < 4096 || A.max_width > 32) return;
    constexpr int kSample = 1024;
    int64_t *d = nullptr;
    dev_alloc(&d, sizeof(int64_t) * kSample * 32);
    k_row_offsets<<<(kSample + kWG - 1) / kWG, kWG, 0, s>>>(A.nrows, kSample, A.slice_ptr,
                                                            A.col, d);
    std::vector<int64_t> h((size_t)kSample * 32);
    HIPCHECK(hipMemcpyAsync(h.data(), d, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    dev_free(d);
    // plane: median over rows of the largest offset
    std::vector<int64_t> m(kSample);
    for (int k = 0; k < kSample; ++k)
        m[k] = *std::max_element(h.begin() + (int64_t)k * 32, h.begin() + (int64_t)(k + 1) * 32);
    std::vector<int64_t> mm = m;
    std::nth_element(mm.begin(), mm.begin() + kSample / 2, mm.end());
    const int64_t reach = mm[kSample / 2];
    A.plane = reach / 64 >= 2 ? reach / 64 : 0;
    // halo: median over rows of the largest |offset| below half the reach
    for (int k = 0; k < kSample; ++k) {
        int64_t b = 0;
        for (int j = 0; j < 32; ++j) {
            const int64_t v = h[(size_t)k * 32 + j];
            const int64_t o = v < 0 ? -v : v;
            if (2 * o < reach && o > b) b = o;
        }
        m[k] = b;
    }
    std::nth_element(m.begin(), m.begin() + kSample / 2, m.end());
    A.halo = m[kSample / 2];
}

void spmm(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy, int ncols,
          hipStream_t s)
{
    if (spmm_stencil(A, X, ldx, Y, ldy, ncols, s)) return;
    launch_spmm<double>(A, A.val, X, ldx, Y, ldy, ncols, s);
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

// Blocks with both sides above 8 go to the matrix cores; thinner ones
// (the one-column products of the column-by-column orthogonalization) to
// the streaming kernel, with X and Y swapped when X is the thin side.
static bool gram_is_thin(int m, int n) { return m <= 8 || n <= 8; }

static int64_t thin_scratch(int64_t nrows, int m, int n)
{
    ThinPlan p = (n <= 8) ? thin_plan(nrows, m, n) : thin_plan(nrows, n, m);
    return (int64_t)p.nchunk * p.mtiles * p.mt * p.nb;
}

GramPlan gram_plan(int64_t nrows, int m, int n, bool lower_only)
{
    GramPlan p;
    if (m <= 0 || n <= 0) return p;
    if (gram_is_thin(m, n)) {
        p.thin = true;
        p.scratch = thin_scratch(nrows, m, n);
        return p;
    }
    p.choice = gram_mfma_choice(nrows, m, n, lower_only, &p.tune);
    p.scratch = gram_mfma_scratch(nrows, m, n, p.choice);
    return p;
}

GramPlan gram_plan_bounded(int64_t nrows, int m, int n, bool lower_only, int64_t max_scratch)
{
    GramPlan p = gram_plan(nrows, m, n, lower_only);
    p.tune = false;
    if (p.scratch <= max_scratch || p.thin) {
        if (p.scratch > max_scratch) p.scratch = -1;
        return p;
    }
    // as many row chunks as fit
    const int64_t per = gram_mfma_chunk_doubles(m, n);
    const int64_t fit = max_scratch / per;
    p.choice = -1;
    p.max_chunks = fit;
    p.scratch = fit < 1 ? -1 : gram_mfma_scratch(nrows, m, n, -1, fit);
    return p;
}

void gram(const GramPlan &p, int64_t nrows, int m, int n, const double *X, int64_t ldx,
          const double *Y, int64_t ldy, double *G, int ldg, bool lower_only, double *partial,
          hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    if (!p.thin) {
        gram_mfma(nrows, m, n, X, ldx, Y, ldy, G, ldg, lower_only, p.choice, p.max_chunks, p.tune,
                  partial, s);
        return;
    }
    // lower_only needs no special case here: the full block is computed
    if (n <= 8)
        gram_thin(nrows, m, n, X, ldx, Y, ldy, G, ldg, false, partial, s);
    else
        gram_thin(nrows, n, m, Y, ldy, X, ldx, G, ldg, true, partial, s);
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

// Single-precision BlockPCG on the device, installed in the table's
// MultiLinearSolver slot (the user-defined solver of
// src/ops_eig_sol_gcg.c:585-642).
//
// It solves (A + sigma B) x_c = b_c for the columns of a block with the
// stopping rule of BlockPCG (src/ops_lin_sol.c:383-394): a column stops
// once its residual is below rate * initial residual or below tol, or
// after max_iter iterations; a stopped column's iterate is frozen.
//
// One call is one step of mixed-precision iterative refinement: the
// residual r = b - op(x) is formed in FP64 (k_resid64), the correction d
// solves op(d) = r in FP32 from d = 0, and x += d in FP64 (k_add_up).  GCG
// tolerates the inexact inner solve; its FP64 residual check and
// Rayleigh-Ritz step refine further.
//
// The FP32 work vectors (d, r, p, w) are private, so they use a layout made
// for the stencil gather: columns in tiles of 8, interleaved by row,
// element (row, col) at ((col / 8) * n + row) * 8 + col % 8.  A neighbour
// row's 8 values are one contiguous 32-byte read per lane (2 KB per
// wavefront).  Per iteration, over the tiles that still hold an active
// column (their list is compacted on the device):
//   k_update_p     d += alpha' p (the previous iteration's step), p = r + beta p
//   k_spmm_pw      w = (A + sigma B) p, partial p.w per workgroup
//   k_finish_pw    alpha = rho / p.w (0 for stopped columns)
//   k_update_xr    r -= alpha w, partial r.r per workgroup
//   k_sum_cols     the r.r sum of each active column (a workgroup per column)
//   k_finish_rr    rho, stop flags, compacted tile list and the tiles it dropped
// and a 4-byte read of the active tile count by the host every kCheck
// iterations.  Reductions are written per workgroup and summed in a fixed
// order (no atomics).  Deferring d += alpha p to the next k_update_p, which
// reads p anyway, saves one pass over p per iteration (8 vector passes
// instead of 9); the tiles a finish drops get their last step there too,
// and the last iteration's step is applied after the loop.  Each d
// receives the same steps in the same order as without the deferral.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "app_hip_internal.h"
#include "gcge_hip.h"
#include "kernels.h"
#include "sell_rows.h"
#include "tuner.h"

namespace gcge {
namespace {

constexpr int kWG = 256;
constexpr int kC = 8;      // columns per interleaved tile
constexpr int kCheck = 4;  // inner iterations between reads of the active count

struct alignas(32) f8 {
    float v[kC];
};

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup of kC per-thread values; result in red[4*kC + t].
__device__ __forceinline__ void block_sum8(const float (&v)[kC], float *red)
{
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < kC; ++t) {
        const float s = wave_sum(v[t]);
        if (ln == 0) red[wv * kC + t] = s;
    }
    __syncthreads();
    if (threadIdx.x < kC) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < kWG / 64; ++k) s += red[k * kC + threadIdx.x];
        red[kWG / 64 * kC + threadIdx.x] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ f8 load8(const float *p)
{
    return *reinterpret_cast<const f8 *>(p);
}

__device__ __forceinline__ void store8(float *p, const f8 &x)
{
    *reinterpret_cast<f8 *>(p) = x;
}

typedef float f4v __attribute__((ext_vector_type(4)));

// load8 through a buffer resource at a byte offset (two 16-byte loads).
// The loaded vectors are bit-cast whole: a bit cast of one element of the
// integer vector reads the vector's first element for every index.
__device__ __forceinline__ f8 buffer_load8(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    const f4v a = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    const f4v b = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0));
    f8 x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        x.v[i] = a[i];
        x.v[4 + i] = b[i];
    }
    return x;
}

// Residual of the FP64 system, rounded to FP32, for every column c < nc:
// R[c] = fl32(b[:, c] - (A + sigma B) x[:, c]) with the products in FP64
// from the FP64 matrix values; partial[c * nblk + block] = sum of R^2.
// Columns nc .. 8*ntile-1 of the last tile are set to zero.  This is the
// refinement step: the FP32 solve then finds the correction, whose size
// is that of the residual, not the whole solution, whose FP32 rounding
// would swamp the correction.
template <int W, bool SHIFT, bool XC>
__global__ __launch_bounds__(kWG) void k_resid64(
    int64_t n, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const double *__restrict__ valA, const int *__restrict__ colB,
    const double *__restrict__ valB, double sigma, int nc, const double *__restrict__ Bv,
    int64_t ldb, const double *__restrict__ Xv, int64_t ldx, float *__restrict__ R,
    float *__restrict__ partial)
{
    static_assert(kWG == 256, "one wavefront per slice, four per workgroup");
    __shared__ float red[(kWG / 64 + 1) * kC];
    const int64_t nblk = gridDim.x;
    // partial sums are stored by row block, so their order (and the
    // rounding of their sum) does not depend on XC
    const int64_t rb = row_block<XC>();
    const int64_t row = rb * kWG + threadIdx.x;
    const bool live = row < n;
    // the wavefront's slice width; every lane of a slice has stored entries
    // (padding for rows past n), so all lanes gather and only live ones
    // store
    const int64_t s = wave_slice<kWG>(rb);
    const int lane = (int)(threadIdx.x & 63);
    int w = 0;
    int64_t base = 0;
    if (s < (n + 63) / 64) {  // wave-uniform
        base = slice_ptr[s];
        w = __builtin_amdgcn_readfirstlane((int)((slice_ptr[s + 1] - base) >> 6));
    }
    uint32_t oa[W], ob[W];
    double va[W], vb[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {  // wave-uniform
            const int64_t e = base + (int64_t)j * 64 + lane;
            oa[j] = (uint32_t)colA[e] * 8u;
            va[j] = valA[e];
            ob[j] = SHIFT ? (uint32_t)colB[e] * 8u : oa[j];
            vb[j] = SHIFT ? sigma * valB[e] : 0.0;
        }
    }
    const uint32_t colbytes = (uint32_t)(n * 8);
    const int ntile = (nc + kC - 1) / kC;
    for (int tile = 0; tile < ntile; ++tile) {
        float sq[kC];
        f8 r8;
#pragma unroll
        for (int t = 0; t < kC; ++t) {
            sq[t] = 0.f;
            r8.v[t] = 0.f;
        }
        for_width<W>(w, [&](auto kw) {
            constexpr int K = decltype(kw)::value;
#pragma unroll
            for (int t = 0; t < kC; ++t) {
                const int c = tile * kC + t;
                if (c < nc) {  // wave-uniform
                    const auto xr = buffer_rsrc(Xv + (int64_t)c * ldx, colbytes);
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        acc += va[j] * buffer_load<double>(xr, oa[j]);
                        if (SHIFT) acc += vb[j] * buffer_load<double>(xr, ob[j]);
                    }
                    if (live) {
                        const float r = (float)(Bv[row + (int64_t)c * ldb] - acc);
                        r8.v[t] = r;
                        sq[t] = r * r;
                    }
                }
            }
        });
        if (live) store8(R + ((int64_t)tile * n + row) * kC, r8);
        block_sum8(sq, red);
        if (threadIdx.x < kC)
            partial[(int64_t)(tile * kC + threadIdx.x) * nblk + rb] =
                red[kWG / 64 * kC + threadIdx.x];
        __syncthreads();
    }
}

// x[:, c] += d[c] in FP64, for c < nc: one thread per row of one tile
// reads the row's 8 values of d at once (a workgroup per column would read
// each 32-byte row segment of d eight times) and updates the tile's columns
// of x, each store coalesced across the wavefront
__global__ __launch_bounds__(kWG) void k_add_up(int64_t n, int nc, const float *__restrict__ D,
                                                double *__restrict__ Xv, int64_t ldx)
{
    const int tile = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r >= n) return;
    const f8 d = load8(D + ((int64_t)tile * n + r) * kC);
    const int c0 = tile * kC;
#pragma unroll
    for (int t = 0; t < kC; ++t)
        if (c0 + t < nc) Xv[r + (int64_t)(c0 + t) * ldx] += (double)d.v[t];
}

// zero the tiles of d
__global__ __launch_bounds__(kWG) void k_zero_tiles(int64_t n, float *__restrict__ D)
{
    const int tile = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r < n) {
        f8 z;
#pragma unroll
        for (int t = 0; t < kC; ++t) z.v[t] = 0.f;
        store8(D + ((int64_t)tile * n + r) * kC, z);
    }
}

// Shifted SpMM on the listed tiles with the p.w partial sums fused:
// w = (A + sigma B) p for the 8 columns of tile tiles[k], k < ntact;
// partial[(tile*8 + t) * nblk + block] = the block's sum of p * w.
// TWO: B is a matrix of its own (a shift by a B other than A); otherwise
// the operator is A (+ sigma A) and only A's entries are held in registers.
// At most 128 VGPRs (four waves per SIMD) except with B's entries gathered
// separately, where the entry offsets and values of both matrices do not
// fit: the unbounded 32-entry kernel took 130 and ran three waves.
template <int W, bool TWO>
__device__ __forceinline__ void spmm_pw_body(int64_t rb, int64_t nblk, 
    int64_t n, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const float *__restrict__ valA, const int *__restrict__ colB,
    const float *__restrict__ valB, float sigma, const int *__restrict__ tiles,
    const int *__restrict__ ntact_p, const float *__restrict__ P, float *__restrict__ Wv,
    float *__restrict__ partial)
{
    static_assert(kWG == 256, "one wavefront per slice, four per workgroup");
    __shared__ float red[(kWG / 64 + 1) * kC];
    const int ntact = *ntact_p;
    const int64_t row = rb * kWG + threadIdx.x;
    const bool live = row < n;
    // the wavefront's slice width; every lane of a slice has stored entries
    // (padding for rows past n), so all lanes gather and only live ones
    // store
    const int64_t s = wave_slice<kWG>(rb);
    const int lane = (int)(threadIdx.x & 63);
    int w = 0;
    int64_t base = 0;
    if (s < (n + 63) / 64) {  // wave-uniform
        base = slice_ptr[s];
        w = __builtin_amdgcn_readfirstlane((int)((slice_ptr[s + 1] - base) >> 6));
    }
    constexpr int WB = TWO ? W : 1;
    uint32_t oa[W], ob[WB];  // byte offsets of the gathered 8-float rows
    float va[W], vb[WB];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {  // wave-uniform
            const int64_t e = base + (int64_t)j * 64 + lane;
            oa[j] = (uint32_t)colA[e] * (uint32_t)(kC * sizeof(float));
            if constexpr (TWO) {
                va[j] = valA[e];
                vb[j] = sigma * valB[e];
                ob[j] = (uint32_t)colB[e] * (uint32_t)(kC * sizeof(float));
            } else {
                va[j] = valB ? valA[e] + sigma * valB[e] : valA[e];
            }
        }
    }
    const uint32_t tbytes = (uint32_t)(n * kC * (int64_t)sizeof(float));
    for (int k = 0; k < ntact; ++k) {
        const int tile = tiles[k];
        const float *pt = P + (int64_t)tile * n * kC;
        const auto pr_rsrc = buffer_rsrc(pt, tbytes);
        float dot[kC];
        f8 acc;
#pragma unroll
        for (int t = 0; t < kC; ++t) {
            acc.v[t] = 0.f;
            dot[t] = 0.f;
        }
        for_width<W>(w, [&](auto kw) {
            constexpr int K = decltype(kw)::value;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const f8 pa = buffer_load8(pr_rsrc, oa[j]);
#pragma unroll
                for (int t = 0; t < kC; ++t) acc.v[t] += va[j] * pa.v[t];
                if constexpr (TWO) {
                    const f8 pb = buffer_load8(pr_rsrc, ob[j]);
#pragma unroll
                    for (int t = 0; t < kC; ++t) acc.v[t] += vb[j] * pb.v[t];
                }
            }
        });
        if (live) {
            store8(Wv + ((int64_t)tile * n + row) * kC, acc);
            const f8 pr = load8(pt + row * kC);
#pragma unroll
            for (int t = 0; t < kC; ++t) dot[t] = acc.v[t] * pr.v[t];
        }
        block_sum8(dot, red);
        if (threadIdx.x < kC)
            partial[(int64_t)(tile * kC + threadIdx.x) * nblk + rb] =
                red[kWG / 64 * kC + threadIdx.x];
        __syncthreads();
    }
}

template <int W, bool TWO, bool XC>
__global__ __launch_bounds__(kWG, (W == 32 && TWO) ? 1 : 4) void k_spmm_pw(
    int64_t n, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const float *__restrict__ valA, const int *__restrict__ colB,
    const float *__restrict__ valB, float sigma, const int *__restrict__ tiles,
    const int *__restrict__ ntact_p, const float *__restrict__ P, float *__restrict__ Wv,
    float *__restrict__ partial)
{
    // partial sums are stored by row block, so their order (and the
    // rounding of their sum) does not depend on XC
    spmm_pw_body<W, TWO>(row_block<XC>(), gridDim.x, n, slice_ptr, colA, valA, colB, valB, sigma,
                         tiles, ntact_p, P, Wv, partial);
}

// p = r + beta p on the listed tiles; p = r on the first iteration (p is
// overwritten without being read: the workspace may hold anything).
// defer: also d += alpha p with the p being replaced and the previous
// iteration's alpha, on the listed tiles and on the `dropped` tiles the last
// finish took off the list (those get only the d step).  only_d: the d step
// alone on both sets (the step of the last iteration, after the loop).
__device__ __forceinline__ void update_p_body(int64_t bx, int64_t gx, int by, 
    int64_t n, const int *__restrict__ tiles, const int *__restrict__ ntact_p,
    const int *__restrict__ dropped, const int *__restrict__ ndropped_p,
    const float *__restrict__ rho, const float *__restrict__ rho_old,
    const float *__restrict__ alpha, int first, int defer, int only_d,
    const float *__restrict__ R, float *__restrict__ P, float *__restrict__ D)
{
    const int k = by;
    const int na = *ntact_p;
    int tile;
    bool step_p;
    if (k < na) {
        tile = tiles[k];
        step_p = !only_d;
    } else if (defer && !first && k < na + *ndropped_p) {
        tile = dropped[k - na];
        step_p = false;
    } else {
        return;
    }
    const bool step_d = defer && !first;
    float beta[kC], a[kC];
#pragma unroll
    for (int t = 0; t < kC; ++t) {
        const int c = tile * kC + t;
        beta[t] = (first || rho_old[c] == 0.f) ? 0.f : rho[c] / rho_old[c];
        a[t] = step_d ? alpha[c] : 0.f;
    }
    const int64_t o = (int64_t)tile * n * kC;
    for (int64_t i = (int64_t)bx * kWG + threadIdx.x; i < n;
         i += (int64_t)gx * kWG) {
        if (first) {
            store8(P + o + i * kC, load8(R + o + i * kC));
            continue;
        }
        f8 p = load8(P + o + i * kC);
        if (step_d) {
            f8 d = load8(D + o + i * kC);
#pragma unroll
            for (int t = 0; t < kC; ++t) d.v[t] += a[t] * p.v[t];
            store8(D + o + i * kC, d);
        }
        if (step_p) {
            const f8 r = load8(R + o + i * kC);
#pragma unroll
            for (int t = 0; t < kC; ++t) p.v[t] = r.v[t] + beta[t] * p.v[t];
            store8(P + o + i * kC, p);
        }
    }
}

__global__ __launch_bounds__(kWG) void k_update_p(
    int64_t n, const int *__restrict__ tiles, const int *__restrict__ ntact_p,
    const int *__restrict__ dropped, const int *__restrict__ ndropped_p,
    const float *__restrict__ rho, const float *__restrict__ rho_old,
    const float *__restrict__ alpha, int first, int defer, int only_d,
    const float *__restrict__ R, float *__restrict__ P, float *__restrict__ D)
{
    update_p_body(blockIdx.x, gridDim.x, blockIdx.y, n, tiles, ntact_p, dropped, ndropped_p, rho,
                  rho_old, alpha, first, defer, only_d, R, P, D);
}

// r -= alpha w on the listed tiles, with r.r partial sums (the d step of
// this iteration is taken by the next k_update_p)
__global__ __launch_bounds__(kWG) void k_update_xr(int64_t n, const int *__restrict__ tiles,
                                                   const int *__restrict__ ntact_p,
                                                   const float *__restrict__ alpha,
                                                   const float *__restrict__ Wv,
                                                   float *__restrict__ R,
                                                   float *__restrict__ partial)
{
    __shared__ float red[(kWG / 64 + 1) * kC];
    const int k = blockIdx.y;
    if (k >= *ntact_p) return;
    const int tile = tiles[k];
    float a[kC], s[kC];
#pragma unroll
    for (int t = 0; t < kC; ++t) {
        a[t] = alpha[tile * kC + t];
        s[t] = 0.f;
    }
    const int64_t o = (int64_t)tile * n * kC;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kWG) {
        const f8 w = load8(Wv + o + i * kC);
        f8 r = load8(R + o + i * kC);
#pragma unroll
        for (int t = 0; t < kC; ++t) {
            r.v[t] -= a[t] * w.v[t];
            s[t] += r.v[t] * r.v[t];
        }
        store8(R + o + i * kC, r);
    }
    block_sum8(s, red);
    if (threadIdx.x < kC)
        partial[(int64_t)(tile * kC + threadIdx.x) * gridDim.x + blockIdx.x] =
            red[kWG / 64 * kC + threadIdx.x];
}

// alpha[c] = rho[c] / (p.w)[c] for the active columns of the listed tiles,
// 0 for stopped ones (their iterate stays frozen).  One wave per column.
__global__ __launch_bounds__(64) void k_finish_pw(const int *__restrict__ tiles,
                                                  const int *__restrict__ ntact_p, int nparts,
                                                  const float *__restrict__ partial,
                                                  const float *__restrict__ rho,
                                                  const int *__restrict__ active,
                                                  float *__restrict__ alpha)
{
    const int k = blockIdx.x / kC, t = blockIdx.x % kC;
    if (k >= *ntact_p) return;
    const int c = tiles[k] * kC + t;
    const float *pc = partial + (int64_t)c * nparts;
    // lane l adds partials l, l + 64, l + 128, ... in that order; the loads
    // of kBatch of them are issued before their adds, so the sum's latency
    // is one load per batch rather than one per partial
    constexpr int kBatch = 16;
    double s = 0.0;
    int i = threadIdx.x;
    for (; i + 64 * (kBatch - 1) < nparts; i += 64 * kBatch) {
        float v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) v[u] = pc[i + 64 * u];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) s += v[u];
    }
    for (; i < nparts; i += 64) s += pc[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x == 0) alpha[c] = active[c] ? (float)((double)rho[c] / s) : 0.f;
}

// Column sums of the r.r partials, one workgroup per column: the workgroup
// copies kSumChunk partials at a time into LDS (coalesced) and one thread
// adds them in index order in FP64, the order a thread walking the column
// would use, while the copies of the next chunk wait at the barrier.
// mode 0: every column c < nc; mode 1: the active ones (the others keep
// their last state and are not read).
constexpr int kSumChunk = 8192;

__global__ __launch_bounds__(256) void k_sum_cols(int nparts, int mode,
                                                  const float *__restrict__ partial,
                                                  const int *__restrict__ active,
                                                  double *__restrict__ sums)
{
    __shared__ float stage[kSumChunk];
    const int c = blockIdx.x;
    if (mode == 1 && active[c] == 0) return;  // workgroup-uniform
    const float *pc = partial + (int64_t)c * nparts;
    double s = 0.0;
    for (int i0 = 0; i0 < nparts; i0 += kSumChunk) {
        const int cnt = nparts - i0 < kSumChunk ? nparts - i0 : kSumChunk;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) stage[i] = pc[i0 + i];
        __syncthreads();
        if (threadIdx.x == 0) {
            int i = 0;
            for (; i + 8 <= cnt; i += 8) {
                float w[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) w[u] = stage[i + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += w[u];
            }
            for (; i < cnt; ++i) s += stage[i];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[c] = s;
}

// New residual norms, stop flags and the compacted tile list from the
// column sums of k_sum_cols; one workgroup, one thread per column
// (<= 1024).  mode 0: initial residual of every column c < nc; mode 1:
// after an iteration, over the active columns.
__global__ __launch_bounds__(1024) void k_finish_rr(int nc, int mode,
                                                    const double *__restrict__ sums,
                                                    int *__restrict__ tiles,
                                                    int *__restrict__ ntact_p,
                                                    int *__restrict__ dropped,
                                                    int *__restrict__ ndropped_p,
                                                    int *__restrict__ active,
                                                    float *__restrict__ rho,
                                                    float *__restrict__ rho_old,
                                                    float *__restrict__ init_res, float rate,
                                                    float tol)
{
    __shared__ int tile_on[1024 / kC], tile_was[1024 / kC];
    const int c = threadIdx.x;
    const int ntile = (nc + kC - 1) / kC;
    if (c < ntile) tile_on[c] = tile_was[c] = 0;
    __syncthreads();
    const bool was = c < nc && (mode == 0 || active[c] != 0);
    // after an iteration the listed tiles are those with an active column
    if (was && mode == 1) tile_was[c / kC] = 1;  // benign race: all writers store 1
    const double s = was ? sums[c] : 0.0;
    if (c < nc) {
        if (was) {
            const float res = (float)std::sqrt(s);
            int on;
            if (mode == 0) {
                init_res[c] = res;
                rho_old[c] = 0.f;
                rho[c] = (float)s;
                on = res > tol ? 1 : 0;
            } else {
                rho_old[c] = rho[c];
                rho[c] = (float)s;
                on = (res > rate * init_res[c] && res > tol) ? 1 : 0;
            }
            active[c] = on;
            if (on) tile_on[c / kC] = 1;  // benign race: all writers store 1
        }
    } else if (c < ntile * kC && mode == 0) {
        // padding columns of the last tile: inactive, with rho = 0 so that
        // beta = 0 and their p, w, d and r stay exactly zero
        active[c] = 0;
        rho[c] = rho_old[c] = init_res[c] = 0.f;
    }
    __syncthreads();
    if (c == 0) {
        int q = 0, qd = 0;
        for (int t = 0; t < ntile; ++t) {
            if (tile_on[t])
                tiles[q++] = t;
            else if (tile_was[t])
                dropped[qd++] = t;
        }
        *ntact_p = q;
        *ndropped_p = qd;
    }
}

// ---------------------------------------------------------------------
// Three launches per iteration for systems of up to kFuseMaxRows rows.
// There the fixed cost of a dependent launch outweighs the work, so the
// two single-workgroup reductions fold into the kernels that follow them:
// each workgroup of k_update_xr_pw forms alpha for its tile's columns from
// the SpMM's partials (the arithmetic of k_finish_pw), and each workgroup
// of k_update_p_rr forms rho and the stop flags of its tile's columns from
// k_update_xr_pw's partials (the arithmetic of k_finish_rr), while its
// workgroup (0, 0) writes the new state of every column and the compacted
// tile list.  That state (rho, stop flags, tile list, count) is double
// buffered by iteration parity, so no kernel overwrites what its own
// workgroups still read; the sums are those of the five-launch iteration.
// Past kFuseMaxRows each workgroup's share of the partials grows with the
// row count and the separate reductions are cheaper.
// ---------------------------------------------------------------------
constexpr int64_t kFuseMaxRows = 65536;

__device__ __forceinline__ void update_xr_pw_body(int64_t bx, int64_t gx, int by, 
    int64_t n, const int *__restrict__ tiles, const int *__restrict__ ntact_p, int nparts_pw,
    const float *__restrict__ partial_pw, const float *__restrict__ rho,
    const int *__restrict__ active, const float *__restrict__ P, const float *__restrict__ Wv,
    float *__restrict__ D, float *__restrict__ R, float *__restrict__ partial_rr)
{
    __shared__ float red[(kWG / 64 + 1) * kC];
    __shared__ float s_alpha[kC];
    const int k = by;
    if (k >= *ntact_p) return;
    const int tile = tiles[k];
    // alpha of the tile's columns, one wavefront per column
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int q = wv; q < kC; q += kWG / 64) {
        const int c = tile * kC + q;
        double sum = 0.0;
        for (int i = ln; i < nparts_pw; i += 64) sum += partial_pw[(int64_t)c * nparts_pw + i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        if (ln == 0) s_alpha[q] = active[c] ? (float)((double)rho[c] / sum) : 0.f;
    }
    __syncthreads();
    float a[kC], sq[kC];
#pragma unroll
    for (int t = 0; t < kC; ++t) {
        a[t] = s_alpha[t];
        sq[t] = 0.f;
    }
    const int64_t o = (int64_t)tile * n * kC;
    for (int64_t i = (int64_t)bx * kWG + threadIdx.x; i < n;
         i += (int64_t)gx * kWG) {
        const f8 p = load8(P + o + i * kC), w = load8(Wv + o + i * kC);
        f8 d = load8(D + o + i * kC), r = load8(R + o + i * kC);
#pragma unroll
        for (int t = 0; t < kC; ++t) {
            d.v[t] += a[t] * p.v[t];
            r.v[t] -= a[t] * w.v[t];
            sq[t] += r.v[t] * r.v[t];
        }
        store8(D + o + i * kC, d);
        store8(R + o + i * kC, r);
    }
    block_sum8(sq, red);
    if (threadIdx.x < kC)
        partial_rr[(int64_t)(tile * kC + threadIdx.x) * gx + bx] =
            red[kWG / 64 * kC + threadIdx.x];
}

__global__ __launch_bounds__(kWG) void k_update_xr_pw(
    int64_t n, const int *__restrict__ tiles, const int *__restrict__ ntact_p, int nparts_pw,
    const float *__restrict__ partial_pw, const float *__restrict__ rho,
    const int *__restrict__ active, const float *__restrict__ P, const float *__restrict__ Wv,
    float *__restrict__ D, float *__restrict__ R, float *__restrict__ partial_rr)
{
    update_xr_pw_body(blockIdx.x, gridDim.x, blockIdx.y, n, tiles, ntact_p, nparts_pw, partial_pw,
                      rho, active, P, Wv, D, R, partial_rr);
}

// rho and stop flag of column c from k_update_xr_pw's partials, as
// k_finish_rr forms them; a stopped column keeps its state.
__device__ __forceinline__ void next_state(int c, int nparts_rr,
                                           const float *__restrict__ partial_rr,
                                           const float *__restrict__ rho_o,
                                           const int *__restrict__ active_o,
                                           const float *__restrict__ init_res, float rate,
                                           float tol, float &rho_n, int &act_n)
{
    rho_n = rho_o[c];
    act_n = active_o[c];
    if (act_n) {
        double sum = 0.0;
        for (int i = 0; i < nparts_rr; ++i) sum += partial_rr[(int64_t)c * nparts_rr + i];
        const float res = (float)std::sqrt(sum);
        rho_n = (float)sum;
        act_n = (res > rate * init_res[c] && res > tol) ? 1 : 0;
    }
}

__device__ __forceinline__ void update_p_rr_body(int64_t bx, int64_t gx, int by, 
    int64_t n, int nc, int nparts_rr, const float *__restrict__ partial_rr,
    const int *__restrict__ tiles_o, const int *__restrict__ ntact_o,
    const float *__restrict__ rho_o, const int *__restrict__ active_o,
    int *__restrict__ tiles_n, int *__restrict__ ntact_n, float *__restrict__ rho_n,
    int *__restrict__ active_n, const float *__restrict__ init_res, float rate, float tol,
    const float *__restrict__ R, float *__restrict__ P)
{
    __shared__ int tile_on[1024 / kC];
    __shared__ float s_beta[kC];
    __shared__ int s_on;
    const int ntile = (nc + kC - 1) / kC;
    if (bx == 0 && by == 0) {
        for (int t = threadIdx.x; t < ntile; t += kWG) tile_on[t] = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < ntile * kC; c += kWG) {
            float r;
            int on;
            next_state(c, nparts_rr, partial_rr, rho_o, active_o, init_res, rate, tol, r, on);
            rho_n[c] = r;
            active_n[c] = on;
            if (on) tile_on[c / kC] = 1;  // benign race: all writers store 1
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int q = 0;
            for (int t = 0; t < ntile; ++t)
                if (tile_on[t]) tiles_n[q++] = t;
            *ntact_n = q;
        }
    }
    const int k = by;
    if (k >= *ntact_o) return;
    const int tile = tiles_o[k];
    if (threadIdx.x == 0) s_on = 0;
    __syncthreads();
    if (threadIdx.x < kC) {
        const int c = tile * kC + threadIdx.x;
        float r;
        int on;
        next_state(c, nparts_rr, partial_rr, rho_o, active_o, init_res, rate, tol, r, on);
        s_beta[threadIdx.x] = rho_o[c] == 0.f ? 0.f : r / rho_o[c];
        if (on) atomicOr(&s_on, 1);
    }
    __syncthreads();
    if (!s_on) return;  // the tile has stopped: it leaves the list
    float beta[kC];
#pragma unroll
    for (int t = 0; t < kC; ++t) beta[t] = s_beta[t];
    const int64_t o = (int64_t)tile * n * kC;
    for (int64_t i = (int64_t)bx * kWG + threadIdx.x; i < n;
         i += (int64_t)gx * kWG) {
        const f8 r = load8(R + o + i * kC);
        f8 p = load8(P + o + i * kC);
#pragma unroll
        for (int t = 0; t < kC; ++t) p.v[t] = r.v[t] + beta[t] * p.v[t];
        store8(P + o + i * kC, p);
    }
}

__global__ __launch_bounds__(kWG) void k_update_p_rr(
    int64_t n, int nc, int nparts_rr, const float *__restrict__ partial_rr,
    const int *__restrict__ tiles_o, const int *__restrict__ ntact_o,
    const float *__restrict__ rho_o, const int *__restrict__ active_o,
    int *__restrict__ tiles_n, int *__restrict__ ntact_n, float *__restrict__ rho_n,
    int *__restrict__ active_n, const float *__restrict__ init_res, float rate, float tol,
    const float *__restrict__ R, float *__restrict__ P)
{
    update_p_rr_body(blockIdx.x, gridDim.x, blockIdx.y, n, nc, nparts_rr, partial_rr, tiles_o,
                     ntact_o, rho_o, active_o, tiles_n, ntact_n, rho_n, active_n, init_res, rate,
                     tol, R, P);
}

// ---------------------------------------------------------------------
// Batched BlockPCG iterations across solves.  A batcher thread (PcgBatcher
// below) runs the three-launch iteration for every solve registered with
// it: one launch per step serves all of them, the solve of workgroup
// (x, y, z) being job z of a table uploaded each round.  Each job runs the
// bodies above with its own launch extents (nb, nblk_row) in place of the
// grid's, its own buffers and its own iteration parity, so its arithmetic
// is that of the same solve run alone, bit for bit.  Grid extents are the
// largest over the jobs; workgroups past their job's extents leave at once.
// ---------------------------------------------------------------------
struct BatchJob {
    int64_t n;
    const int64_t *slice_ptr;
    const int *col;
    const float *val;
    float *P, *Wv, *D, *R, *partial, *partial_rr;
    int *tiles[2], *ntact[2], *active[2];
    float *rho[2];
    const float *init_res;
    float rate, tol;
    int nc, nb, nblk_row, nt;  // nt: bound on the tiles listed in this group of rounds
    int it0, max_iter;         // iterations run before the group; the limit
};

// p = r (first iteration) or p = r + beta p with the state of the previous
// iteration folded in (k_update_p / k_update_p_rr of the fused path)
// Round r of a group runs iteration it0 + r of each job (none past its
// max_iter); its parity picks the state buffers.
__global__ __launch_bounds__(kWG) void k_batch_update_p(const BatchJob *__restrict__ jobs, int r)
{
    const BatchJob &j = jobs[blockIdx.z];
    const int it = j.it0 + r;
    if ((int)blockIdx.x >= j.nb || (int)blockIdx.y >= j.nt || it >= j.max_iter)
        return;  // workgroup-uniform
    const int par = it & 1;
    if (it == 0) {
        update_p_body(blockIdx.x, j.nb, blockIdx.y, j.n, j.tiles[0], j.ntact[0], nullptr, nullptr,
                      j.rho[0], j.rho[0], nullptr, 1, 0, 0, j.R, j.P, j.D);
    } else {
        const int o = par ^ 1, q = par;
        update_p_rr_body(blockIdx.x, j.nb, blockIdx.y, j.n, j.nc, j.nb, j.partial_rr, j.tiles[o],
                         j.ntact[o], j.rho[o], j.active[o], j.tiles[q], j.ntact[q], j.rho[q],
                         j.active[q], j.init_res, j.rate, j.tol, j.R, j.P);
    }
}

// w = A p with the partial p.w sums (k_spmm_pw, rows of at most 8 entries,
// no shift)
__global__ __launch_bounds__(kWG, 4) void k_batch_spmm_pw(const BatchJob *__restrict__ jobs, int r)
{
    const BatchJob &j = jobs[blockIdx.z];
    const int it = j.it0 + r;
    if ((int)blockIdx.x >= j.nblk_row || it >= j.max_iter) return;  // workgroup-uniform
    const int par = it & 1;
    spmm_pw_body<8, false>(blockIdx.x, j.nblk_row, j.n, j.slice_ptr, j.col, j.val, j.col, nullptr,
                           0.f, j.tiles[par], j.ntact[par], j.P, j.Wv, j.partial);
}

// alpha, d += alpha p, r -= alpha w with the r.r partials (k_update_xr_pw)
__global__ __launch_bounds__(kWG) void k_batch_update_xr(const BatchJob *__restrict__ jobs, int r)
{
    const BatchJob &j = jobs[blockIdx.z];
    const int it = j.it0 + r;
    if ((int)blockIdx.x >= j.nb || (int)blockIdx.y >= j.nt || it >= j.max_iter)
        return;  // workgroup-uniform
    const int q = it & 1;
    update_xr_pw_body(blockIdx.x, j.nb, blockIdx.y, j.n, j.tiles[q], j.ntact[q], j.nblk_row,
                      j.partial, j.rho[q], j.active[q], j.P, j.Wv, j.D, j.R, j.partial_rr);
}

// counts[z] = the active-tile count of job z's latest state after a group
// of `rounds` rounds (the state of the iteration before its last one, as
// the solo loop reads it)
__global__ void k_batch_counts(const BatchJob *__restrict__ jobs, int njobs, int rounds,
                               int *__restrict__ counts)
{
    const int z = (int)threadIdx.x;
    if (z < njobs) {
        const BatchJob &j = jobs[z];
        const int last = min(j.it0 + rounds, j.max_iter) - 1;
        counts[z] = *j.ntact[last & 1];
    }
}

struct PCG32 {
    int max_iter = 30;
    double rate = 1e-2, tol = 1e-14;
    void *B = nullptr;
    double sigma = 0.0;
    // device workspace: ncap columns (a multiple of 8) of n rows each
    int64_t n = 0;
    int ncap = 0;
    float *x = nullptr, *r = nullptr, *p = nullptr, *w = nullptr;
    float *rho = nullptr, *rho_old = nullptr, *init_res = nullptr, *alpha = nullptr;
    float *partial = nullptr;
    double *colsum = nullptr;  // per-column sums of the r.r partials
    int *tiles = nullptr, *ntact = nullptr, *active = nullptr;
    int *dropped = nullptr, *ndropped = nullptr;  // tiles the last finish took off the list
    // three-launch iterations: the state of odd iterations, r.r partials
    float *rho_b = nullptr, *partial_rr = nullptr;
    int *tiles_b = nullptr, *ntact_b = nullptr, *active_b = nullptr;
    int *h_ntact = nullptr;
    int64_t iters = 0;
    bool use_fused = true;  // three launches per iteration up to kFuseMaxRows rows
    // the batcher this solver's iterations join (service workers), or null
    struct PcgBatcher *batcher = nullptr;
    // kCheck inner iterations captured as one graph, for one key
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    // everything the captured launches take by value: the launch grids
    // (active count, block width), the matrices, the shift, the stopping
    // rule and the row-block order; the count before the latest read bounds
    // the first k_update_p of a group
    struct GraphKey {
        int ntact = -1, ntact_prev = -1, nc = 0;
        int64_t n = 0;
        const void *a = nullptr, *as = nullptr, *av = nullptr, *b = nullptr, *bv = nullptr;
        float sigma = 0.f, rate = 0.f, tol = 0.f;
        bool xc = false, fused = false;
        bool operator==(const GraphKey &o) const
        {
            return ntact == o.ntact && ntact_prev == o.ntact_prev && nc == o.nc && n == o.n &&
                   a == o.a && as == o.as &&
                   av == o.av && b == o.b && bv == o.bv && sigma == o.sigma && rate == o.rate &&
                   tol == o.tol && xc == o.xc && fused == o.fused;
        }
    } gkey;
    void drop_graph()
    {
        if (gexec) HIPCHECK(hipGraphExecDestroy(gexec));
        gexec = nullptr;
        gkey = GraphKey();
    }
    void ensure(int64_t nrows, int nc)
    {
        const int nc8 = (nc + kC - 1) / kC * kC;
        if (nrows == n && nc8 <= ncap) return;
        release();
        drop_graph();
        n = nrows;
        ncap = nc8;
        const size_t blkbytes = sizeof(float) * (size_t)n * ncap;
        for (float **q : {&x, &r, &p, &w})
            dev_alloc(q, blkbytes);
        for (float **q : {&rho, &rho_old, &init_res, &alpha})
            dev_alloc(q, sizeof(float) * ncap);
        dev_alloc(&partial, sizeof(float) * (size_t)ncap * nparts_max());
        dev_alloc(&colsum, sizeof(double) * ncap);
        dev_alloc(&tiles, sizeof(int) * (ncap / kC));
        dev_alloc(&active, sizeof(int) * ncap);
        dev_alloc(&ntact, sizeof(int));
        dev_alloc(&dropped, sizeof(int) * (ncap / kC));
        dev_alloc(&ndropped, sizeof(int));
        pinned_alloc(&h_ntact, sizeof(int));
        dev_alloc(&rho_b, sizeof(float) * ncap);
        dev_alloc(&active_b, sizeof(int) * ncap);
        dev_alloc(&tiles_b, sizeof(int) * (ncap / kC));
        dev_alloc(&ntact_b, sizeof(int));
        dev_alloc(&partial_rr, sizeof(float) * (size_t)ncap * 1024);
    }
    int64_t nparts_max() const { return (n + kWG - 1) / kWG; }
    void release()
    {
        drop_graph();
        if (!x) return;
        ctx().sync();
        for (float *q : {x, r, p, w, rho, rho_old, init_res, alpha, partial}) dev_free(q);
        dev_free(colsum);
        dev_free(tiles);
        dev_free(active);
        dev_free(ntact);
        dev_free(dropped);
        dev_free(ndropped);
        pinned_free(h_ntact);
        dev_free(rho_b);
        dev_free(active_b);
        dev_free(tiles_b);
        dev_free(ntact_b);
        dev_free(partial_rr);
        x = nullptr;
    }
};

unsigned blocks_for(int64_t n, int cap)
{
    int64_t b = (n + kWG * 4 - 1) / (kWG * 4);
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

// The batcher: a thread that runs rounds of batched iterations over the
// solves submitted to it, kCheck rounds per group.  A solve is submitted
// after its initial residual and state are in place (the submitting thread
// has waited for its stream); it joins the next group, leaves once its
// active-tile count reads 0 or it has run max_iter iterations, gets
// x += d, and its caller is woken.  The job table is uploaded once per
// group and the counts of all the jobs are read after it (one copy each
// way and one wait per group); between reads the kernels skip the tiles
// that have stopped, so rounds after a job's last active iteration change
// nothing.
struct PcgBatcher {
    struct Job {
        BatchJob d;
        int it = 0, max_iter = 0, nt = 0;
        double mat_bytes = 0.0, nnz = 0.0;  // of the job's matrix, for the kernel timing
        double *x64 = nullptr;
        int64_t ldx = 0;
        int64_t *iters = nullptr;
        bool done = false;
    };
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<Job *> incoming;
    bool stop = false;
    std::thread th;
    int64_t rounds = 0, job_iters = 0;
    int max_jobs_seen = 0;

    PcgBatcher() { th = std::thread([this] { run(); }); }
    ~PcgBatcher()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        th.join();
    }

    // runs the job's iterations with the others'; returns when done
    void solve(Job &j)
    {
        std::unique_lock<std::mutex> lk(mu);
        incoming.push_back(&j);
        cv_work.notify_all();
        cv_done.wait(lk, [&] { return j.done; });
    }

    void run()
    {
        constexpr int kMaxJobs = 256;
        Ctx &c = ctx();
        hipStream_t st = c.stream;
        BatchJob *h_tab = nullptr, *d_tab = nullptr;
        int *h_cnt = nullptr, *d_cnt = nullptr;
        pinned_alloc(&h_tab, sizeof(BatchJob) * kMaxJobs);
        dev_alloc(&d_tab, sizeof(BatchJob) * kMaxJobs);
        pinned_alloc(&h_cnt, sizeof(int) * kMaxJobs);
        dev_alloc(&d_cnt, sizeof(int) * kMaxJobs);
        std::vector<Job *> active;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !incoming.empty() || !active.empty(); });
                if (stop && incoming.empty() && active.empty()) break;
                while (!incoming.empty() && (int)active.size() < kMaxJobs) {
                    active.push_back(incoming.back());
                    incoming.pop_back();
                }
            }
            const int nj = (int)active.size();
            if (nj > max_jobs_seen) max_jobs_seen = nj;
            int gx = 1, gy = 1, gr = 1, nr = 1;
            for (int z = 0; z < nj; ++z) {
                Job &j = *active[z];
                j.d.it0 = j.it;
                j.d.max_iter = j.max_iter;
                j.d.nt = j.nt;
                h_tab[z] = j.d;
                gx = std::max(gx, j.d.nb);
                gy = std::max(gy, j.nt);
                gr = std::max(gr, j.d.nblk_row);
                nr = std::max(nr, std::min(kCheck, j.max_iter - j.it));
            }
            HIPCHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(BatchJob) * nj, hipMemcpyHostToDevice, st));
            for (int r = 0; r < nr; ++r) {
                // algorithmic bytes and flops of the round, summed over the
                // jobs it advances (the solo loop's formulas over each job's
                // listed tiles)
                double bp = 0.0, fp = 0.0, bs = 0.0, fs = 0.0, bx = 0.0, fx = 0.0;
                for (int z = 0; z < nj; ++z) {
                    const Job &j = *active[z];
                    if (j.it + r >= j.max_iter) continue;
                    const double tv = 4.0 * kC * (double)j.d.n * j.nt;
                    bp += (j.it + r == 0 ? 2.0 : 3.0) * tv;
                    fp += 2.0 * kC * (double)j.d.n * j.nt;
                    bs += j.mat_bytes + 2.0 * tv;
                    fs += 2.0 * j.nnz * kC * j.nt;
                    bx += 6.0 * tv;
                    fx += 6.0 * kC * (double)j.d.n * j.nt;
                }
                hipEvent_t e1 = c.prof.begin(PROF_PCGVEC, bp, fp, st);
                k_batch_update_p<<<dim3(gx, gy, nj), kWG, 0, st>>>(d_tab, r);
                c.prof.end(e1, st);
                hipEvent_t e2 = c.prof.begin(PROF_SPMM32, bs, fs, st);
                k_batch_spmm_pw<<<dim3(gr, 1, nj), kWG, 0, st>>>(d_tab, r);
                c.prof.end(e2, st);
                hipEvent_t e3 = c.prof.begin(PROF_PCGVEC, bx, fx, st);
                k_batch_update_xr<<<dim3(gx, gy, nj), kWG, 0, st>>>(d_tab, r);
                c.prof.end(e3, st);
            }
            k_batch_counts<<<1, kMaxJobs, 0, st>>>(d_tab, nj, nr, d_cnt);
            HIPCHECK(hipMemcpyAsync(h_cnt, d_cnt, sizeof(int) * nj, hipMemcpyDeviceToHost, st));
            c.sync();
            rounds += nr;
            std::vector<Job *> finished;
            for (int z = 0; z < nj; ++z) {
                Job &j = *active[z];
                const int ran = std::min(nr, j.max_iter - j.it);
                job_iters += ran;
                j.it += ran;
                j.nt = h_cnt[z];
                if (j.nt == 0 || j.it >= j.max_iter) finished.push_back(&j);
            }
            if (finished.empty()) continue;
            for (Job *j : finished) {
                const unsigned nblk = (unsigned)((j->d.n + kWG - 1) / kWG);
                const int ntile = (j->d.nc + kC - 1) / kC;
                k_add_up<<<dim3(nblk, ntile), kWG, 0, st>>>(j->d.n, j->d.nc, j->d.D, j->x64,
                                                            j->ldx);
                *j->iters += j->it;
            }
            c.sync();
            // a Job lives in its caller's frame, which is gone once the
            // caller wakes: leave the finished ones out of the list before
            // they are released and never touch them afterwards
            active.erase(std::remove_if(active.begin(), active.end(),
                                        [&](Job *j) {
                                            return std::find(finished.begin(), finished.end(),
                                                             j) != finished.end();
                                        }),
                         active.end());
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Job *j : finished) j->done = true;
            }
            cv_done.notify_all();
        }
        pinned_free(h_tab);
        dev_free(d_tab);
        pinned_free(h_cnt);
        dev_free(d_cnt);
        release_ctx();
    }
};

void pcg32_solve(void *matA, void **mv_b, void **mv_x, int *start, int *end, OPS *ops)
{
    PCG32 &S = *static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    Mat *A = mat_of(matA, "BlockPCG32 A");
    MV *Bv = mv_of(mv_b, "BlockPCG32 b"), *Xv = mv_of(mv_x, "BlockPCG32 x");
    const int nc = end[0] - start[0];
    if (nc != end[1] - start[1]) fail("BlockPCG32: ranges differ", __FILE__, __LINE__);
    if (nc <= 0) return;
    if (nc > 1024) fail("BlockPCG32: more than 1024 columns", __FILE__, __LINE__);
    if (A->sell.val32 == nullptr) fail("BlockPCG32: A has no FP32 copy", __FILE__, __LINE__);
    if (A->sell.max_width > 32)
        fail("BlockPCG32: rows wider than 32 entries are not supported", __FILE__, __LINE__);
    const SellMat *Bs = &A->sell;
    float sigma = 0.f;
    if (S.B != nullptr && S.sigma != 0.0) {
        Mat *Bm = mat_of(S.B, "BlockPCG32 B");
        if (Bm->sell.val32 == nullptr) fail("BlockPCG32: B has no FP32 copy", __FILE__, __LINE__);
        // the shifted kernels walk A's slice pointers for both matrices
        if (Bm->sell.structure_hash != A->sell.structure_hash)
            fail("BlockPCG32: a shift needs A and B with the same SELL row structure", __FILE__,
                 __LINE__);
        Bs = &Bm->sell;
        sigma = (float)S.sigma;
    }
    const int64_t n = A->n;
    // gathers use 32-bit byte offsets into one 8-column tile
    if (n * kC * (int64_t)sizeof(float) > (int64_t)UINT32_MAX)
        fail("BlockPCG32: more than 2^27 rows", __FILE__, __LINE__);
    S.ensure(n, nc);
    Ctx &c = ctx();
    hipStream_t st = c.stream;
    const int ntile = (nc + kC - 1) / kC;
    const unsigned nblk_row = (unsigned)((n + kWG - 1) / kWG);
    // B's entries enter separately only for a shift by a B other than A
    const bool two = sigma != 0.f && Bs->col != A->sell.col;
    // row-block order of the gathering kernels (XCD-contiguous or plain):
    // the table's choice, tuned below on the first iteration's operands
    const TuneKey pkey{kTuneSpmmPw, (A->sell.max_width <= 8 ? 8 : 32) + (two ? 100 : 0),
                       ilog2_floor(n), 0};
    int pch = tune_lookup(pkey);
    const bool tune_pw = (pch < 0 || pch > 1) && tune_enabled();
    if (pch < 0 || pch > 1) pch = 0;
    bool xc = pch == 1;
    k_zero_tiles<<<dim3(nblk_row, ntile), kWG, 0, st>>>(n, S.x);
    {
        const double *b64 = Bv->d + (int64_t)start[0] * Bv->ldd;
        const double *x64 = Xv->d + (int64_t)start[1] * Xv->ldd;
        const double *vB = (Bs != &A->sell) ? Bs->val : nullptr;
#define RESID(W, SH, XC)                                                                      \
    k_resid64<W, SH, XC><<<nblk_row, kWG, 0, st>>>(n, A->sell.slice_ptr, A->sell.col,            \
                                                   A->sell.val, Bs->col, vB, S.sigma, nc, b64,   \
                                                   Bv->ldd, x64, Xv->ldd, S.r, S.partial)
        const bool sh = vB != nullptr;
        if (A->sell.max_width <= 8) {
            if (sh) {
                if (xc) RESID(8, true, true); else RESID(8, true, false);
            } else {
                if (xc) RESID(8, false, true); else RESID(8, false, false);
            }
        } else {
            if (sh) {
                if (xc) RESID(32, true, true); else RESID(32, true, false);
            } else {
                if (xc) RESID(32, false, true); else RESID(32, false, false);
            }
        }
#undef RESID
    }
    k_sum_cols<<<nc, 256, 0, st>>>((int)nblk_row, 0, S.partial, S.active, S.colsum);
    k_finish_rr<<<1, 1024, 0, st>>>(nc, 0, S.colsum, S.tiles, S.ntact, S.dropped, S.ndropped,
                                    S.active, S.rho, S.rho_old, S.init_res, (float)S.rate,
                                    (float)S.tol);
    HIPCHECK(hipMemcpyAsync(S.h_ntact, S.ntact, sizeof(int), hipMemcpyDeviceToHost, st));
    c.sync();
    int ntact = *S.h_ntact;
    const unsigned nb = blocks_for(n, 1024);
    // algorithmic bytes of one shifted SpMM over the active tiles: the CCS
    // arrays of A (and of B when shifted) once, p read once, w written once
    // A constant-coefficient 7-point A without shift is applied matrix-free
    // (stencil.hip), with its own count of p.w partial sums per column
    // (its most partial sums per column: those of a launch over one tile)
    const int sp = sigma == 0.f ? stencil_pw_parts(A->sell, 1) : 0;
    const int sparts = sp <= S.nparts_max() ? sp : 0;
    // partial sums per column of a product launched over at most nt tiles
    auto nparts_pw = [&](int nt) {
        return sparts > 0 ? stencil_pw_parts(A->sell, nt) : (int)nblk_row;
    };
    // served solves of the fused kind (general rows of at most 8 entries,
    // no shift) run their iterations in the batcher's rounds
    if (S.batcher != nullptr && S.use_fused && n <= kFuseMaxRows && sparts == 0 && !two &&
        sigma == 0.f && A->sell.max_width <= 8) {
        PcgBatcher::Job j;
        BatchJob &d = j.d;
        d.n = n;
        d.slice_ptr = A->sell.slice_ptr;
        d.col = A->sell.col;
        d.val = A->sell.val32;
        d.P = S.p;
        d.Wv = S.w;
        d.D = S.x;
        d.R = S.r;
        d.partial = S.partial;
        d.partial_rr = S.partial_rr;
        d.tiles[0] = S.tiles;
        d.tiles[1] = S.tiles_b;
        d.ntact[0] = S.ntact;
        d.ntact[1] = S.ntact_b;
        d.active[0] = S.active;
        d.active[1] = S.active_b;
        d.rho[0] = S.rho;
        d.rho[1] = S.rho_b;
        d.init_res = S.init_res;
        d.rate = (float)S.rate;
        d.tol = (float)S.tol;
        d.nc = nc;
        d.nb = (int)nb;
        d.nblk_row = (int)nblk_row;
        j.max_iter = S.max_iter;
        j.nt = ntact;
        j.mat_bytes = (double)A->sell.nnz_true * 8.0 + 4.0 * (double)(n + 1);
        j.nnz = (double)A->sell.nnz_true;
        j.x64 = Xv->d + (int64_t)start[1] * Xv->ldd;
        j.ldx = Xv->ldd;
        j.iters = &S.iters;
        if (ntact > 0 && S.max_iter > 0) {
            S.batcher->solve(j);
        } else {
            k_add_up<<<dim3(nblk_row, ntile), kWG, 0, st>>>(n, nc, S.x, j.x64, j.ldx);
        }
        return;
    }
    const double mat_bytes =
        sparts > 0 ? 0.0
                   : (double)A->sell.nnz_true * 8.0 + 4.0 * (double)(n + 1) +
                         (sigma != 0.f ? (double)Bs->nnz_true * 8.0 : 0.0);
    int it = 0;
    // algorithmic bytes and flops of the vector updates over the active
    // tiles (as of the host's last read of the count, which can only
    // overstate them: exact while no tile has stopped).  Five launches:
    // d += alpha' p, p = r + beta p reads p, d, r and writes p, d (r read
    // and p written on the first iteration); r -= alpha w (with r.r) reads
    // w, r and writes r.  Three launches: p = r + beta p reads r, p and
    // writes p; d += alpha p, r -= alpha w reads p, w, d, r and writes d, r.
    const double tile_vec = 4.0 * kC * (double)n;
    const float *vB = sigma != 0.f ? Bs->val32 : nullptr;
    auto spmm_pw = [&](bool x, const int *tl, const int *nt_p, int nt_bound) {
        if (sparts > 0) {
            stencil_pw(A->sell, tl, nt_p, nt_bound, S.p, S.w, S.partial, st);
            return;
        }
#define SPMM_PW(W, TWO, XC)                                                                   \
    k_spmm_pw<W, TWO, XC><<<nblk_row, kWG, 0, st>>>(n, A->sell.slice_ptr, A->sell.col,        \
                                                    A->sell.val32, Bs->col, vB, sigma, tl,     \
                                                    nt_p, S.p, S.w, S.partial)
        if (A->sell.max_width <= 8) {
            if (two) {
                if (x) SPMM_PW(8, true, true); else SPMM_PW(8, true, false);
            } else {
                if (x) SPMM_PW(8, false, true); else SPMM_PW(8, false, false);
            }
        } else {
            if (two) {
                if (x) SPMM_PW(32, true, true); else SPMM_PW(32, true, false);
            } else {
                if (x) SPMM_PW(32, false, true); else SPMM_PW(32, false, false);
            }
        }
#undef SPMM_PW
    };
    if (tune_pw && ntact > 0 && sparts == 0) {
        // the tile list and count are set; p is written by the first
        // iteration before it is read, and w and the partials are
        // overwritten by every launch, so the timing runs leave no trace
        const int best =
            tune_run(pkey, 2, 0, [&](int i) { spmm_pw(i == 1, S.tiles, S.ntact, ntact); }, st);
        xc = best == 1;
    }
    // three launches per iteration up to kFuseMaxRows rows, with the state
    // of iteration i in buffer i % 2 (the state the first finish wrote is
    // buffer 0); five past it
    const bool fused = S.use_fused && n <= kFuseMaxRows;
    int *tl_buf[2] = {S.tiles, S.tiles_b}, *nt_buf[2] = {S.ntact, S.ntact_b};
    int *act_buf[2] = {S.active, S.active_b};
    float *rho_buf[2] = {S.rho, S.rho_b};
    // one inner iteration on st; `nt` bounds the tiles listed for it, `nt_p`
    // those listed for the previous one (the five-launch k_update_p also
    // takes the d step of the tiles the latest finish dropped); `prof`
    // records the kernel-class events; `par` is the iteration's parity
    auto iteration = [&](int first, int nt, int nt_p, bool prof, int par) {
        const double pass_p = first ? 2.0 : (fused ? 3.0 : 5.0);
        hipEvent_t evp = prof ? c.prof.begin(PROF_PCGVEC, pass_p * tile_vec * nt,
                                             (fused ? 2.0 : 4.0) * kC * (double)n * nt, st)
                              : nullptr;
        if (!fused || first) {
            // no reduction in this kernel: one row per thread (measured 4-5 %
            // faster than grid-stride loops over 1024 workgroups per tile)
            k_update_p<<<dim3(fused ? nb : nblk_row, fused ? nt : nt_p), kWG, 0, st>>>(
                n, S.tiles, S.ntact, S.dropped, S.ndropped, S.rho, S.rho_old, S.alpha, first,
                fused ? 0 : 1, 0, S.r, S.p, S.x);
        } else {
            const int o = par ^ 1;
            k_update_p_rr<<<dim3(nb, nt), kWG, 0, st>>>(
                n, nc, (int)nb, S.partial_rr, tl_buf[o], nt_buf[o], rho_buf[o], act_buf[o],
                tl_buf[par], nt_buf[par], rho_buf[par], act_buf[par], S.init_res,
                (float)S.rate, (float)S.tol, S.r, S.p);
        }
        if (prof) c.prof.end(evp, st);
        const int *tl = fused ? tl_buf[par] : S.tiles;
        const int *ntp = fused ? nt_buf[par] : S.ntact;
        hipEvent_t ev = prof ? c.prof.begin(PROF_SPMM32, mat_bytes + 2.0 * 4.0 * kC * (double)nt * n,
                                            2.0 * (double)A->sell.nnz_true * kC * nt, st)
                             : nullptr;
        spmm_pw(xc, tl, ntp, nt);
        if (prof) c.prof.end(ev, st);
        hipEvent_t evx = prof ? c.prof.begin(PROF_PCGVEC, (fused ? 6.0 : 3.0) * tile_vec * nt,
                                             (fused ? 6.0 : 4.0) * kC * (double)n * nt, st)
                              : nullptr;
        if (fused) {
            k_update_xr_pw<<<dim3(nb, nt), kWG, 0, st>>>(
                n, tl, ntp, nparts_pw(nt), S.partial, rho_buf[par], act_buf[par], S.p, S.w, S.x,
                S.r, S.partial_rr);
            if (prof) c.prof.end(evx, st);
        } else {
            k_finish_pw<<<nt * kC, 64, 0, st>>>(S.tiles, S.ntact, nparts_pw(nt), S.partial,
                                                S.rho, S.active, S.alpha);
            k_update_xr<<<dim3(nb, nt), kWG, 0, st>>>(n, S.tiles, S.ntact, S.alpha, S.w, S.r,
                                                      S.partial);
            if (prof) c.prof.end(evx, st);
            k_sum_cols<<<nc, 256, 0, st>>>((int)nb, 1, S.partial, S.active, S.colsum);
            k_finish_rr<<<1, 1024, 0, st>>>(nc, 1, S.colsum, S.tiles, S.ntact, S.dropped,
                                            S.ndropped, S.active, S.rho, S.rho_old, S.init_res,
                                            (float)S.rate, (float)S.tol);
        }
    };
    int last_nt = ntact;  // the tile count the latest iteration launched with
    // the host's count before its latest read: an iteration right after a
    // read takes the d step of tiles listed for the previous one, which the
    // new count may not cover
    int prev_nt = ntact;
    while (it < S.max_iter && ntact > 0) {
        // the host reads the active-tile count every kCheck iterations; in
        // between it launches with the last count it read, an upper bound
        // (tiles only ever stop), and the kernels skip the tiles beyond the
        // device-side count.  A whole group of kCheck iterations after the
        // first is one graph launch (launch overhead dominates small
        // problems), unless kernel timing is on.  A graph is captured only
        // while this thread's context is the process's only one; one
        // captured earlier is replayed regardless.
        const bool whole = it > 0 && it % kCheck == 0 && it + kCheck <= S.max_iter;
        PCG32::GraphKey key;
        key.ntact = ntact;
        key.nc = nc;
        key.n = n;
        key.a = A->sell.col;
        key.as = A->sell.slice_ptr;
        key.av = A->sell.val32;
        key.b = Bs->col;
        key.bv = vB;
        key.sigma = sigma;
        key.rate = (float)S.rate;
        key.tol = (float)S.tol;
        key.xc = xc;
        key.fused = fused;
        key.ntact_prev = fused ? -1 : prev_nt;  // only five-launch groups use it
        const bool have = S.gexec && key == S.gkey;
        std::unique_lock<std::mutex> solo;
        if (whole && S.use_graphs && !c.prof.on && !have) solo = sole_context_lock();
        if (whole && S.use_graphs && !c.prof.on && (have || solo.owns_lock())) {
            if (!have) {
                S.drop_graph();
                hipGraph_t graph;
                HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                // groups start at multiples of kCheck (even): parities 0, 1,
                // 0, 1; the first iteration follows a read of the count
                for (int k = 0; k < kCheck; ++k)
                    iteration(0, ntact, k == 0 ? prev_nt : ntact, false, k % 2);
                HIPCHECK(hipStreamEndCapture(st, &graph));
                HIPCHECK(hipGraphInstantiate(&S.gexec, graph, nullptr, nullptr, 0));
                HIPCHECK(hipGraphDestroy(graph));
                S.gkey = key;
                solo.unlock();
            }
            HIPCHECK(hipGraphLaunch(S.gexec, st));
            it += kCheck;
        } else {
            iteration(it == 0 ? 1 : 0, ntact, it % kCheck == 0 ? prev_nt : ntact, true, it % 2);
            ++it;
        }
        last_nt = ntact;
        if (it % kCheck == 0 || it == S.max_iter) {
            // the latest state: with three launches, that of iteration it - 1
            // (which counts the stops of iteration it - 2; a tile that stopped
            // in the last iteration leaves the list in the next one)
            const int *ntp = fused ? nt_buf[(it - 1) % 2] : S.ntact;
            HIPCHECK(hipMemcpyAsync(S.h_ntact, ntp, sizeof(int), hipMemcpyDeviceToHost, st));
            c.sync();
            prev_nt = ntact;
            ntact = *S.h_ntact;
        }
    }
    S.iters += it;
    if (!fused && it > 0) {
        // the last iteration's d step, on the tiles it ran on: the list of
        // its finish and the tiles that finish dropped (last_nt, the count
        // its launches used, bounds both)
        hipEvent_t evf = c.prof.begin(PROF_PCGVEC, 3.0 * tile_vec * last_nt,
                                      2.0 * kC * (double)n * last_nt, st);
        k_update_p<<<dim3(nblk_row, last_nt), kWG, 0, st>>>(n, S.tiles, S.ntact, S.dropped,
                                                      S.ndropped, S.rho, S.rho_old, S.alpha, 0,
                                                      1, 1, S.r, S.p, S.x);
        c.prof.end(evf, st);
    }
    // x += d in FP64
    k_add_up<<<dim3(nblk_row, ntile), kWG, 0, st>>>(n, nc, S.x,
                                                    Xv->d + (int64_t)start[1] * Xv->ldd, Xv->ldd);
}

}  // namespace

void *pcg32_batcher_create() { return new PcgBatcher(); }

void pcg32_batcher_destroy(void *b) { delete static_cast<PcgBatcher *>(b); }

void pcg32_batcher_stats(void *b, int64_t *rounds, int64_t *job_iters, int *max_jobs)
{
    PcgBatcher *p = static_cast<PcgBatcher *>(b);
    std::lock_guard<std::mutex> lk(p->mu);
    if (rounds) *rounds = p->rounds;
    if (job_iters) *job_iters = p->job_iters;
    if (max_jobs) *max_jobs = p->max_jobs_seen;
}

void pcg32_set_batcher(OPS *ops, void *b)
{
    if (ops->MultiLinearSolver != pcg32_solve || ops->multi_linear_solver_workspace == nullptr)
        fail("pcg32_set_batcher: no device BlockPCG installed", __FILE__, __LINE__);
    static_cast<PCG32 *>(ops->multi_linear_solver_workspace)->batcher =
        static_cast<PcgBatcher *>(b);
}

int64_t pcg32_iterations(OPS *ops)
{
    if (ops->MultiLinearSolver != pcg32_solve || ops->multi_linear_solver_workspace == nullptr)
        return 0;
    return static_cast<const PCG32 *>(ops->multi_linear_solver_workspace)->iters;
}

}  // namespace gcge

extern "C" void HIPGCG_MultiLinearSolverFree_BlockPCG32(OPS *ops)
{
    using namespace gcge;
    if (ops->MultiLinearSolver != pcg32_solve || ops->multi_linear_solver_workspace == nullptr)
        return;
    PCG32 *S = static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    S->release();
    delete S;
    ops->multi_linear_solver_workspace = nullptr;
    ops->MultiLinearSolver = nullptr;
}

extern "C" void HIPGCG_BlockPCG32_SetFused(OPS *ops, int on)
{
    using namespace gcge;
    if (ops->MultiLinearSolver != pcg32_solve || ops->multi_linear_solver_workspace == nullptr)
        fail("HIPGCG_BlockPCG32_SetFused: no device BlockPCG installed", __FILE__, __LINE__);
    PCG32 *S = static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    S->use_fused = on != 0;
    S->drop_graph();
}

extern "C" void HIPGCG_BlockPCG32_SetGraphs(OPS *ops, int on)
{
    using namespace gcge;
    if (ops->MultiLinearSolver != pcg32_solve || ops->multi_linear_solver_workspace == nullptr)
        fail("HIPGCG_BlockPCG32_SetGraphs: no device BlockPCG installed", __FILE__, __LINE__);
    PCG32 *S = static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    S->use_graphs = on != 0;
    if (!S->use_graphs) S->drop_graph();
}

extern "C" void HIPGCG_MultiLinearSolverSetup_BlockPCG32(int max_iter, double rate, double tol,
                                                         void *B, double sigma, OPS *ops)
{
    using namespace gcge;
    PCG32 *S = static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    if (S == nullptr || ops->MultiLinearSolver != pcg32_solve) S = new PCG32();
    S->drop_graph();  // captured with the previous parameters
    S->max_iter = max_iter;
    S->rate = rate;
    S->tol = tol;
    S->B = B;
    S->sigma = sigma;
    ops->multi_linear_solver_workspace = S;
    ops->MultiLinearSolver = pcg32_solve;
}

// Single-precision BlockPCG on the device, installed in the table's
// MultiLinearSolver slot (the user-defined solver of
// src/ops_eig_sol_gcg.c:585-642).
//
// It solves (A + sigma B) x_c = b_c for the columns of a block with the
// stopping rule of BlockPCG (src/ops_lin_sol.c:383-394): a column stops
// once its residual is below rate * initial residual or below tol, or
// after max_iter iterations.
//
// One call is one step of mixed-precision iterative refinement: the
// residual r = b - op(x) is formed in FP64 (k_resid64), the correction d
// solves op(d) = r in FP32 from d = 0, and x += d in FP64 (k_add_up).  GCG
// tolerates the inexact inner solve; its FP64 residual check and
// Rayleigh-Ritz step refine further.
//
// Per iteration, for the active columns only (their indices are compacted
// on the device):
//   k_update_p     p = r + beta p                       (beta = rho/rho_old)
//   k_spmm_pw      w = (A + sigma B) p, partial p.w per workgroup
//   k_finish_pw    alpha = rho / p.w
//   k_update_xr    x += alpha p, r -= alpha w, partial r.r per workgroup
//   k_finish_rr    rho_new, convergence flags, compacted active list
// and one 4-byte read of the active count by the host.  All reductions are
// written per workgroup and summed in a fixed order (no atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "app_hip_internal.h"
#include "gcge_hip.h"
#include "kernels.h"

namespace gcge {
namespace {

constexpr int kWG = 256;
constexpr int kXCD = 8;
constexpr int kT = 8;  // columns per register tile

__device__ __forceinline__ int64_t xcd_swizzle(int64_t bid, int64_t nblk)
{
    const int64_t q = nblk / kXCD, r = nblk % kXCD;
    const int64_t x = bid % kXCD, k = bid / kXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup of kT per-thread values; results in red[0..kT).
__device__ __forceinline__ void block_sum_tile(float (&v)[kT], float *red /* [4][kT] */)
{
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < kT; ++t) {
        const float s = wave_sum(v[t]);
        if (ln == 0) red[wv * kT + t] = s;
    }
    __syncthreads();
    if (threadIdx.x < kT) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < kWG / 64; ++k) s += red[k * kT + threadIdx.x];
        red[kWG / 64 * kT + threadIdx.x] = s;
    }
    __syncthreads();
}

// Residual of the FP64 system, rounded to FP32: for the columns c < nc,
// R32[:, c] = fl32(b[:, c] - (A + sigma B) x[:, c]) with the products in
// FP64 from the FP64 matrix values, and partial[c * nblk + block] = the
// block's sum of R32^2.  This is the refinement step: the FP32 solve then
// finds the correction, whose size is that of the residual, instead of
// the whole solution, whose FP32 rounding would swamp the correction.
template <int W>
__global__ __launch_bounds__(kWG) void k_resid64(
    int64_t n, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const double *__restrict__ valA, const int *__restrict__ colB,
    const double *__restrict__ valB, double sigma, int nc, const double *__restrict__ Bv,
    int64_t ldb, const double *__restrict__ Xv, int64_t ldx, float *__restrict__ R32,
    int64_t ldr, float *__restrict__ partial)
{
    __shared__ float red[(kWG / 64 + 1) * kT];
    const int64_t nblk = gridDim.x;
    const int64_t blk = xcd_swizzle(blockIdx.x, nblk);
    const int64_t row = blk * kWG + threadIdx.x;
    const bool live = row < n;
    const bool shifted = (sigma != 0.0) && (valB != nullptr);
    int ca[W], cb[W];
    double va[W], vb[W];
    int w = 0;
    if (live) {
        const int64_t s = row >> 6;
        const int lane = (int)(row & 63);
        const int64_t base = slice_ptr[s];
        w = (int)((slice_ptr[s + 1] - base) >> 6);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (j < w) {
                const int64_t e = base + (int64_t)j * 64 + lane;
                ca[j] = colA[e];
                va[j] = valA[e];
                cb[j] = shifted ? colB[e] : ca[j];
                vb[j] = shifted ? sigma * valB[e] : 0.0;
            } else {
                ca[j] = cb[j] = (int)row;
                va[j] = vb[j] = 0.0;
            }
        }
    }
    for (int c0 = 0; c0 < nc; c0 += kT) {
        float sq[kT];
#pragma unroll
        for (int t = 0; t < kT; ++t) sq[t] = 0.f;
        if (live) {
#pragma unroll
            for (int t = 0; t < kT; ++t) {
                if (c0 + t < nc) {
                    const double *xc = Xv + (int64_t)(c0 + t) * ldx;
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < W; ++j)
                        if (j < w) acc += va[j] * xc[ca[j]] + vb[j] * xc[cb[j]];
                    const float r = (float)(Bv[row + (int64_t)(c0 + t) * ldb] - acc);
                    R32[row + (int64_t)(c0 + t) * ldr] = r;
                    sq[t] = r * r;
                }
            }
        }
        block_sum_tile(sq, red);
        if (threadIdx.x < kT && c0 + (int)threadIdx.x < nc)
            partial[(int64_t)(c0 + threadIdx.x) * nblk + blockIdx.x] =
                red[kWG / 64 * kT + threadIdx.x];
        __syncthreads();
    }
}

// x[:, c] += (double) d[:, c]  and  d[:, c] = 0 (ready for the next solve)
__global__ __launch_bounds__(kWG) void k_add_up(int64_t n, float *__restrict__ D, int64_t ldd,
                                                double *__restrict__ Xv, int64_t ldx)
{
    const int c = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r < n) {
        Xv[r + (int64_t)c * ldx] += (double)D[r + (int64_t)c * ldd];
        D[r + (int64_t)c * ldd] = 0.f;
    }
}

__global__ __launch_bounds__(kWG) void k_zero32(int64_t n, float *__restrict__ D, int64_t ldd)
{
    const int c = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (r < n) D[r + (int64_t)c * ldd] = 0.f;
}

// Shifted SpMM on the listed columns with the p.w partial sums fused.
// w[:, c] = (A + sigma B) p[:, c] for c = cols[k], k < nact;
// partial[k * nblk + block] = sum over the block's rows of p * w.
// When with_dot == 0 the partial sums are skipped (initial residual).
template <int W>
__global__ __launch_bounds__(kWG) void k_spmm_pw(
    int64_t n, const int64_t *__restrict__ slice_ptr, const int *__restrict__ colA,
    const float *__restrict__ valA, const int *__restrict__ colB, const float *__restrict__ valB,
    float sigma, const int *__restrict__ cols, const int *__restrict__ nact_p,
    const float *__restrict__ P, float *__restrict__ Wv, int64_t ld, float *__restrict__ partial,
    int with_dot)
{
    __shared__ float red[(kWG / 64 + 1) * kT];
    const int nact = *nact_p;
    const int64_t nblk = gridDim.x;
    const int64_t blk = xcd_swizzle(blockIdx.x, nblk);
    const int64_t row = blk * kWG + threadIdx.x;
    const bool live = row < n;
    int ca[W], cb[W];
    float va[W], vb[W];
    int w = 0;
    const bool same = (colA == colB);
    if (live) {
        const int64_t s = row >> 6;
        const int lane = (int)(row & 63);
        const int64_t base = slice_ptr[s];
        w = (int)((slice_ptr[s + 1] - base) >> 6);
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
                va[j] = vb[j] = 0.f;
            }
        }
    }
    for (int k0 = 0; k0 < nact; k0 += kT) {
        float acc[kT], dot[kT];
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            acc[t] = 0.f;
            dot[t] = 0.f;
        }
        if (live) {
            int64_t off[kT];
#pragma unroll
            for (int t = 0; t < kT; ++t)
                off[t] = (k0 + t < nact) ? (int64_t)cols[k0 + t] * ld : -1;
#pragma unroll
            for (int j = 0; j < W; ++j) {
                if (j < w) {
#pragma unroll
                    for (int t = 0; t < kT; ++t) {
                        if (off[t] >= 0) {
                            const float *pc = P + off[t];
                            if (same)
                                acc[t] += (va[j] + vb[j]) * pc[ca[j]];
                            else
                                acc[t] += va[j] * pc[ca[j]] + vb[j] * pc[cb[j]];
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < kT; ++t) {
                if (off[t] >= 0) {
                    Wv[off[t] + row] = acc[t];
                    if (with_dot) dot[t] = acc[t] * P[off[t] + row];
                }
            }
        }
        if (with_dot) {
            block_sum_tile(dot, red);
            if (threadIdx.x < kT && k0 + (int)threadIdx.x < nact)
                partial[(int64_t)(k0 + threadIdx.x) * nblk + blockIdx.x] =
                    red[kWG / 64 * kT + threadIdx.x];
            __syncthreads();
        }
    }
}

// p = r + beta p on the active columns; beta = 0 on the first iteration.
__global__ __launch_bounds__(kWG) void k_update_p(int64_t n, const int *__restrict__ cols,
                                                  const int *__restrict__ nact_p,
                                                  const float *__restrict__ rho,
                                                  const float *__restrict__ rho_old, int first,
                                                  const float *__restrict__ R,
                                                  float *__restrict__ P, int64_t ld)
{
    const int k = blockIdx.y;
    if (k >= *nact_p) return;
    const int c = cols[k];
    const float beta = first ? 0.f : rho[c] / rho_old[c];
    const float *r = R + (int64_t)c * ld;
    float *p = P + (int64_t)c * ld;
    // the first iteration overwrites p without reading it: the workspace
    // may hold anything, including NaN
    if (first) {
        for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kWG)
            p[i] = r[i];
    } else {
        for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kWG)
            p[i] = r[i] + beta * p[i];
    }
}

// x += alpha p, r -= alpha w on the active columns, with r.r partial sums.
__global__ __launch_bounds__(kWG) void k_update_xr(int64_t n, const int *__restrict__ cols,
                                                   const int *__restrict__ nact_p,
                                                   const float *__restrict__ alpha,
                                                   const float *__restrict__ P,
                                                   const float *__restrict__ Wv,
                                                   float *__restrict__ X, float *__restrict__ R,
                                                   int64_t ld, float *__restrict__ partial)
{
    __shared__ float red[kWG / 64];
    const int k = blockIdx.y;
    if (k >= *nact_p) return;
    const int c = cols[k];
    const float a = alpha[c];
    const int64_t o = (int64_t)c * ld;
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kWG) {
        X[o + i] += a * P[o + i];
        const float rv = R[o + i] - a * Wv[o + i];
        R[o + i] = rv;
        s += rv * rv;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < kWG / 64; ++q) t += red[q];
        partial[(int64_t)k * gridDim.x + blockIdx.x] = t;
    }
}

// alpha[c] = rho[c] / sum(partial) for the active columns (one wave each)
__global__ __launch_bounds__(64) void k_finish_pw(const int *__restrict__ cols,
                                                  const int *__restrict__ nact_p, int nparts,
                                                  const float *__restrict__ partial,
                                                  const float *__restrict__ rho,
                                                  float *__restrict__ alpha)
{
    const int k = blockIdx.x;
    if (k >= *nact_p) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) s += partial[(int64_t)k * nparts + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x == 0) {
        const int c = cols[k];
        alpha[c] = (float)((double)rho[c] / s);
    }
}

// New residual norms, convergence and the compacted active list; a single
// workgroup handles every column (<= 1024).  mode 0: initial residual
// (all columns listed in order), mode 1: iteration.
__global__ __launch_bounds__(1024) void k_finish_rr(int ncols, int mode, int nparts,
                                                    const float *__restrict__ partial,
                                                    int *__restrict__ cols,
                                                    int *__restrict__ nact_p,
                                                    float *__restrict__ rho,
                                                    float *__restrict__ rho_old,
                                                    float *__restrict__ init_res, float rate,
                                                    float tol)
{
    __shared__ int keep[1024];
    __shared__ int pos[1024];
    const int k = threadIdx.x;
    const int nact = (mode == 0) ? ncols : *nact_p;
    int c = -1, kp = 0;
    if (k < nact) {
        c = (mode == 0) ? k : cols[k];
        double s = 0.0;
        for (int i = 0; i < nparts; ++i) s += partial[(int64_t)k * nparts + i];
        const float res = (float)std::sqrt(s);
        if (mode == 0) {
            init_res[c] = res;
            rho[c] = (float)s;
            kp = res > tol ? 1 : 0;
        } else {
            rho_old[c] = rho[c];
            rho[c] = (float)s;
            kp = (res > rate * init_res[c] && res > tol) ? 1 : 0;
        }
    }
    keep[k] = kp;
    __syncthreads();
    if (k == 0) {
        int q = 0;
        for (int i = 0; i < nact; ++i) {
            pos[i] = q;
            q += keep[i];
        }
        *nact_p = q;
    }
    __syncthreads();
    if (k < nact && kp) cols[pos[k]] = c;
}

struct PCG32 {
    int max_iter = 30;
    double rate = 1e-2, tol = 1e-14;
    void *B = nullptr;
    double sigma = 0.0;
    // device workspace, FP32, ncap columns of n rows each
    int64_t n = 0;
    int ncap = 0;
    float *x = nullptr, *r = nullptr, *p = nullptr, *w = nullptr;
    float *rho = nullptr, *rho_old = nullptr, *init_res = nullptr, *alpha = nullptr;
    float *partial = nullptr;
    int *cols = nullptr, *nact = nullptr;
    int *h_nact = nullptr;
    int64_t iters = 0;
    void ensure(int64_t nrows, int nc)
    {
        if (nrows == n && nc <= ncap) return;
        release();
        n = nrows;
        ncap = nc;
        const size_t blkbytes = sizeof(float) * (size_t)n * nc;
        for (float **q : {&x, &r, &p, &w}) HIPCHECK(hipMalloc(reinterpret_cast<void **>(q), blkbytes));
        for (float **q : {&rho, &rho_old, &init_res, &alpha})
            HIPCHECK(hipMalloc(reinterpret_cast<void **>(q), sizeof(float) * nc));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&partial),
                           sizeof(float) * (size_t)nc * nparts_max()));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&cols), sizeof(int) * nc));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&nact), sizeof(int)));
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&h_nact), sizeof(int), hipHostMallocDefault));
    }
    int64_t nparts_max() const { return (n + kWG - 1) / kWG; }
    void release()
    {
        if (!x) return;
        ctx().sync();
        for (float *q : {x, r, p, w, rho, rho_old, init_res, alpha, partial}) HIPCHECK(hipFree(q));
        HIPCHECK(hipFree(cols));
        HIPCHECK(hipFree(nact));
        HIPCHECK(hipHostFree(h_nact));
        x = nullptr;
    }
};

unsigned blocks_for(int64_t n, int cap)
{
    int64_t b = (n + kWG * 4 - 1) / (kWG * 4);
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

void spmm_pw_launch(const SellMat &A, const SellMat &Bm, float sigma, const int *cols,
                    const int *nact, const float *P, float *Wv, int64_t ld, float *partial,
                    int with_dot, hipStream_t s)
{
    const unsigned g = (unsigned)((A.nrows + kWG - 1) / kWG);
    if (A.max_width <= 8)
        k_spmm_pw<8><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, A.val32, Bm.col, Bm.val32,
                                       sigma, cols, nact, P, Wv, ld, partial, with_dot);
    else if (A.max_width <= 32)
        k_spmm_pw<32><<<g, kWG, 0, s>>>(A.nrows, A.slice_ptr, A.col, A.val32, Bm.col, Bm.val32,
                                        sigma, cols, nact, P, Wv, ld, partial, with_dot);
    else
        fail("BlockPCG32: rows wider than 32 entries are not supported", __FILE__, __LINE__);
}

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
    const SellMat *Bs = &A->sell;
    float sigma = 0.f;
    if (S.B != nullptr && S.sigma != 0.0) {
        Mat *Bm = mat_of(S.B, "BlockPCG32 B");
        if (Bm->sell.val32 == nullptr) fail("BlockPCG32: B has no FP32 copy", __FILE__, __LINE__);
        if (Bm->sell.nslices != A->sell.nslices)
            fail("BlockPCG32: A and B differ in size", __FILE__, __LINE__);
        Bs = &Bm->sell;
        sigma = (float)S.sigma;
    }
    // the shifted kernels walk A's slice pointers for both matrices
    if (Bs != &A->sell && Bs->structure_hash != A->sell.structure_hash)
        fail("BlockPCG32: a shift needs A and B with the same SELL row structure", __FILE__,
             __LINE__);
    const int64_t n = A->n;
    S.ensure(n, nc);
    Ctx &c = ctx();
    hipStream_t st = c.stream;
    const dim3 gcol((unsigned)((n + kWG - 1) / kWG), nc);
    // refinement: r = fl32(b - op(x)) in FP64, solve op(d) = r from d = 0
    k_zero32<<<gcol, kWG, 0, st>>>(n, S.x, n);
    const unsigned nblk_row = (unsigned)((n + kWG - 1) / kWG);
    {
        const double *b64 = Bv->d + (int64_t)start[0] * Bv->ldd;
        const double *x64 = Xv->d + (int64_t)start[1] * Xv->ldd;
        const double *vB = (Bs != &A->sell) ? Bs->val : nullptr;
        if (A->sell.max_width <= 8)
            k_resid64<8><<<nblk_row, kWG, 0, st>>>(n, A->sell.slice_ptr, A->sell.col, A->sell.val,
                                                   Bs->col, vB, S.sigma, nc, b64, Bv->ldd, x64,
                                                   Xv->ldd, S.r, n, S.partial);
        else if (A->sell.max_width <= 32)
            k_resid64<32><<<nblk_row, kWG, 0, st>>>(n, A->sell.slice_ptr, A->sell.col,
                                                    A->sell.val, Bs->col, vB, S.sigma, nc, b64,
                                                    Bv->ldd, x64, Xv->ldd, S.r, n, S.partial);
        else
            fail("BlockPCG32: rows wider than 32 entries are not supported", __FILE__, __LINE__);
    }
    {
        std::vector<int> all((size_t)nc);
        for (int k = 0; k < nc; ++k) all[k] = k;
        HIPCHECK(hipMemcpyAsync(S.cols, all.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
    }
    const unsigned nb = blocks_for(n, 1024);
    k_finish_rr<<<1, 1024, 0, st>>>(nc, 0, (int)nblk_row, S.partial, S.cols, S.nact, S.rho,
                                    S.rho_old, S.init_res, (float)S.rate, (float)S.tol);
    HIPCHECK(hipMemcpyAsync(S.h_nact, S.nact, sizeof(int), hipMemcpyDeviceToHost, st));
    c.sync();
    int nact = *S.h_nact;
    const int nparts_spmm = (int)((n + kWG - 1) / kWG);
    // algorithmic bytes of one shifted SpMM over k columns: the CCS arrays
    // of A (and of B when shifted) once, p read once, w written once
    const double mat_bytes = (double)A->sell.nnz_true * 8.0 + 4.0 * (double)(n + 1) +
                             (sigma != 0.f ? (double)Bs->nnz_true * 8.0 : 0.0);
    int it = 0;
    while (it < S.max_iter && nact > 0) {
        k_update_p<<<dim3(nb, nact), kWG, 0, st>>>(n, S.cols, S.nact, S.rho, S.rho_old,
                                                   it == 0 ? 1 : 0, S.r, S.p, n);
        hipEvent_t ev = c.prof.begin(PROF_SPMM32, mat_bytes + 2.0 * 4.0 * (double)nact * n,
                                     2.0 * (double)A->sell.nnz_true * nact, st);
        spmm_pw_launch(A->sell, *Bs, sigma, S.cols, S.nact, S.p, S.w, n, S.partial, 1, st);
        c.prof.end(ev, st);
        k_finish_pw<<<nact, 64, 0, st>>>(S.cols, S.nact, nparts_spmm, S.partial, S.rho, S.alpha);
        k_update_xr<<<dim3(nb, nact), kWG, 0, st>>>(n, S.cols, S.nact, S.alpha, S.p, S.w, S.x,
                                                    S.r, n, S.partial);
        k_finish_rr<<<1, 1024, 0, st>>>(nc, 1, (int)nb, S.partial, S.cols, S.nact, S.rho,
                                        S.rho_old, S.init_res, (float)S.rate, (float)S.tol);
        HIPCHECK(hipMemcpyAsync(S.h_nact, S.nact, sizeof(int), hipMemcpyDeviceToHost, st));
        c.sync();
        nact = *S.h_nact;
        ++it;
    }
    S.iters += it;
    // x += d in FP64 (and d back to zero)
    k_add_up<<<gcol, kWG, 0, st>>>(n, S.x, n, Xv->d + (int64_t)start[1] * Xv->ldd, Xv->ldd);
}

}  // namespace
}  // namespace gcge

extern "C" void HIPGCG_MultiLinearSolverSetup_BlockPCG32(int max_iter, double rate, double tol,
                                                         void *B, double sigma, OPS *ops)
{
    using namespace gcge;
    PCG32 *S = static_cast<PCG32 *>(ops->multi_linear_solver_workspace);
    if (S == nullptr || ops->MultiLinearSolver != pcg32_solve) S = new PCG32();
    S->max_iter = max_iter;
    S->rate = rate;
    S->tol = tol;
    S->B = B;
    S->sigma = sigma;
    ops->multi_linear_solver_workspace = S;
    ops->MultiLinearSolver = pcg32_solve;
}

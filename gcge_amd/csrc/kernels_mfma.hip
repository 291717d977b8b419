// FP64 matrix-core (MFMA) kernels for the dense contractions of GCG:
// the Gram blocks of MultiVecInnerProd / MultiVecQtAP (X^T Y) and the
// tall-skinny products of MultiVecLinearComb (X C).
//
// v_mfma_f64_16x16x4_f64 computes D(16x16) += A(16x4) B(4x16) on one
// wavefront.  Lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4v][l%16] in element v of its 4-double accumulator (a layout
// any slip in which fails the Gram and LinearComb parity tests of
// tests/test_gpu_slots.py).
#include <hip/hip_runtime.h>
#include <cstdint>

#include "app_hip_internal.h"
#include "kernels.h"
#include "tuner.h"

namespace gcge {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kWG = 256;
constexpr int kXCD = 8;

__device__ __forceinline__ int64_t xcd_swizzle(int64_t bid, int64_t nblk)
{
    const int64_t q = nblk / kXCD, r = nblk % kXCD;
    const int64_t x = bid % kXCD, k = bid / kXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------
// Gram: G (m x n) = X^T Y.  A workgroup owns a 128 x 128 tile of G over one
// chunk of rows; its four waves own 64 x 64 quadrants (4 x 4 MFMA tiles).
// Rows are staged 16 at a time through LDS, stored column-major (row index
// fastest, padded to 17) so each 16-lane MFMA operand read covers the
// banks once.  Each thread fetches 4 consecutive rows (one 32-byte load)
// of 4 columns per operand, and the next stage's loads are issued into
// registers before the current stage's MFMAs, so HBM latency overlaps the
// matrix-core work; at 98 VGPRs + 128 AGPRs two waves share each SIMD.
// Partial tiles per chunk are summed by k_gram_mfma_reduce in a fixed
// order (deterministic, no atomics).
// ---------------------------------------------------------------------
constexpr int kGT = 128;
constexpr int kGK = 16;                 // rows per stage
constexpr int kGKP = kGK + 1;
constexpr int kGRG = kGK / 4;           // 4-row groups per stage
constexpr int kGCT = kGT * kGRG / kWG;  // columns per thread per operand

__global__ __launch_bounds__(kWG) void k_gram_mfma(
    int64_t nrows, int m, int n, const double *__restrict__ X, int64_t ldx,
    const double *__restrict__ Y, int64_t ldy, int tm, int tn, int64_t rows_per_chunk,
    int lower_only, double *__restrict__ partial)
{
    __shared__ double sx[kGT * kGKP];
    __shared__ double sy[kGT * kGKP];
    const int64_t lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tiles = tm * tn;
    const int chunk = (int)(lb / tiles);
    const int t = (int)(lb % tiles);
    const int bi = t / tn, bj = t % tn;
    double *out = partial + ((int64_t)chunk * tiles + t) * (kGT * kGT);
    if (lower_only && bi < bj) {
        for (int e = threadIdx.x; e < kGT * kGT; e += kWG) out[e] = 0.0;
        return;
    }
    const int i0 = bi * kGT, j0 = bj * kGT;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t rbeg = (int64_t)chunk * rows_per_chunk;
    int64_t rend = rbeg + rows_per_chunk;
    if (rend > nrows) rend = nrows;
    // this thread's slots in a stage: rows 4*rg .. 4*rg+3 of columns
    // kGCT*cg .. kGCT*cg+kGCT-1 (one 32-byte load per column; kGRG lanes
    // cover the stage's rows of a column, so a wave instruction reads whole
    // row segments of 64/kGRG columns, and only kGCT column addresses per
    // operand stay live)
    const int rg = threadIdx.x % kGRG;
    const int cg = threadIdx.x / kGRG;
    const double *xc[kGCT], *yc[kGCT];
    bool xok[kGCT], yok[kGCT];
#pragma unroll
    for (int c = 0; c < kGCT; ++c) {
        const int cc = cg * kGCT + c;
        xok[c] = i0 + cc < m;
        yok[c] = j0 + cc < n;
        xc[c] = X + (int64_t)(xok[c] ? i0 + cc : 0) * ldx + 4 * rg;
        yc[c] = Y + (int64_t)(yok[c] ? j0 + cc : 0) * ldy + 4 * rg;
    }
    double px[kGCT][4], py[kGCT][4];
    auto fetch = [&](int64_t r0) {
        const int64_t r = r0 + 4 * rg;
        if (r + 4 <= rend) {
#pragma unroll
            for (int c = 0; c < kGCT; ++c) {
                const d4 vx = xok[c] ? *reinterpret_cast<const d4 *>(xc[c] + r0)
                                     : d4{0.0, 0.0, 0.0, 0.0};
                const d4 vy = yok[c] ? *reinterpret_cast<const d4 *>(yc[c] + r0)
                                     : d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    px[c][t] = vx[t];
                    py[c][t] = vy[t];
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < kGCT; ++c)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool rok = r + t < rend;
                    px[c][t] = (rok && xok[c]) ? xc[c][r0 + t] : 0.0;
                    py[c][t] = (rok && yok[c]) ? yc[c][r0 + t] : 0.0;
                }
        }
    };
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    if (rbeg < rend) fetch(rbeg);
    for (int64_t r0 = rbeg; r0 < rend; r0 += kGK) {
#pragma unroll
        for (int c = 0; c < kGCT; ++c)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sx[(cg * kGCT + c) * kGKP + 4 * rg + t] = px[c][t];
                sy[(cg * kGCT + c) * kGKP + 4 * rg + t] = py[c][t];
            }
        __syncthreads();
        if (r0 + kGK < rend) fetch(r0 + kGK);
#pragma unroll 2
        for (int kk = 0; kk < kGK / 4; ++kk) {
            const int row = kk * 4 + lk;
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sx[(wm * 64 + a * 16 + li) * kGKP + row];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = sy[(wn * 64 + b * 16 + li) * kGKP + row];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = mfma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = wm * 64 + a * 16 + lk + 4 * v;
                const int jj = wn * 64 + b * 16 + li;
                out[ii * kGT + jj] = acc[a][b][v];
            }
}

__global__ __launch_bounds__(kWG) void k_gram_mfma_reduce(int m, int n, int tm, int tn,
                                                          int nchunk,
                                                          const double *__restrict__ partial,
                                                          double *__restrict__ G, int ldg)
{
    // threads walk the partial tiles in storage order, so each chunk's read
    // is contiguous across the wavefront
    const int64_t off = (int64_t)blockIdx.x * kWG + threadIdx.x;
    const int64_t stride = (int64_t)tm * tn * kGT * kGT;
    if (off >= stride) return;
    const int t = (int)(off / (kGT * kGT));
    const int within = (int)(off % (kGT * kGT));
    const int i = (t / tn) * kGT + within / kGT, j = (t % tn) * kGT + within % kGT;
    if (i >= m || j >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[off + (int64_t)c * stride];
    G[i + (int64_t)j * ldg] = s;
}

struct GramPlanM {
    int tm, tn, nchunk;
    int64_t rows_per_chunk;
};

// Row chunking: at most max_wg workgroups in total (256 CUs, two resident
// each, at the default 512: one more than that runs alone in a second
// round and doubles the time) and at least min_rows rows (min_rows / 16
// stages) per chunk, so small problems still spread over many workgroups
// instead of a few walking all their rows.
GramPlanM plan_with(int64_t nrows, int m, int n, int64_t min_rows, int64_t max_wg)
{
    GramPlanM p;
    p.tm = (m + kGT - 1) / kGT;
    p.tn = (n + kGT - 1) / kGT;
    const int64_t tiles = (int64_t)p.tm * p.tn;
    int64_t want = max_wg / tiles;
    int64_t cap = (nrows + min_rows - 1) / min_rows;
    int64_t nc = want < cap ? want : cap;
    if (nc < 1) nc = 1;
    int64_t rpc = (nrows + nc - 1) / nc;
    rpc = (rpc + kGK - 1) / kGK * kGK;
    p.rows_per_chunk = rpc;
    p.nchunk = (int)((nrows + rpc - 1) / rpc);
    if (p.nchunk < 1) p.nchunk = 1;
    return p;
}

// Tuner candidates (min_rows, max_wg); the default is index kGramDefault.
constexpr int64_t kGramMinRows[5] = {64, 128, 256, 512, 1024};
constexpr int64_t kGramMaxWg[2] = {512, 1024};
constexpr int kGramCands = kGramMfmaCands;
constexpr int kGramDefault = 3 * 2 + 0;  // 512 rows, 512 workgroups

GramPlanM plan_cand(int64_t nrows, int m, int n, int i)
{
    return plan_with(nrows, m, n, kGramMinRows[i / 2], kGramMaxWg[i % 2]);
}

TuneKey gram_key(int64_t nrows, int m, int n, bool lower_only)
{
    const int tiles = ((m + kGT - 1) / kGT) * ((n + kGT - 1) / kGT);
    return TuneKey{kTuneGram, ilog2_floor(nrows), tiles < 64 ? tiles : 64, lower_only ? 1 : 0};
}


// ---------------------------------------------------------------------
// LinearComb: Y (nrows x n) = X (nrows x m) C (m x n) + Y diag(beta),
// evaluated as Y^T = C^T X^T.  A wave owns 64 rows x 64 columns of Y
// (4 x 4 MFMA tiles): per step of 4 k it takes four X operands (lane l
// reads X[i + l%16, k + l/16], 128-byte row segments straight from HBM)
// and four C^T operands (from L1/L2: C is small) and issues 16 MFMAs, so
// each operand feeds four matrix-core instructions.  Operands sit in a
// ring of kRing one-step buffers: the loads for step s + kRing - 1 are
// issued before the MFMAs of step s, and at 224-246 registers two waves share
// each SIMD, so HBM latency hides behind the other wave's MFMAs.  The
// accumulator holds 16 consecutive rows of one column per lane group, so
// the stores are contiguous.  Ct is C^T packed k-major: Ct[k * ldct + j].
// A workgroup (4 waves) covers 256 rows; consecutive workgroups on one XCD
// take the column tiles of the same rows.
// ---------------------------------------------------------------------
// Ring depth by row count: over 2 M rows a 5-step ring ties with 4 without
// beta and is 2-8 % faster with it (m = 220-460); up to 65,536 rows, where
// launches are latency-bound, 5 steps are 1-15 % slower, so smaller
// problems keep 4.
constexpr int64_t kRingDeepRows = int64_t(1) << 18;

template <int BM, int kRing>
__global__ __launch_bounds__(kWG) void k_lincomb_mfma(
    int64_t nrows, int m, int n, const double *__restrict__ X, int64_t ldx,
    const double *__restrict__ Ct, int ldct, double beta_scalar,
    const double *__restrict__ beta, double *__restrict__ Y, int64_t ldy, int ntj)
{
    const int64_t lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tj = (int)(lb % ntj);
    const int64_t tb = lb / ntj;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t i0 = tb * 256 + wave * 64;
    const int j0 = tj * 64;
    bool rok[4];
    const double *xr[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t row = i0 + a * 16 + li;
        rok[a] = row < nrows;
        xr[a] = X + (rok[a] ? row : 0);
    }
    bool jok[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) jok[q] = (j0 + q * 16 + li) < n;
    // BM (beta mode) 0: overwrite; 1: scalar beta; 2: per-column beta.  The
    // accumulators start from Y diag(beta), read 16 at a time (a scheduling
    // barrier after each row tile keeps the 64 loads from being in flight at
    // once, which would double the registers and halve the occupancy)
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        // out-of-range rows and columns read a clamped in-range element and
        // are zeroed by the select (no branch per element)
        const int64_t rowc = rok[a] ? i0 + a * 16 + li : nrows - 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            d4 init;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + q * 16 + lk + 4 * v;
                if (BM == 0) {
                    init[v] = 0.0;
                } else {
                    const int jc = j < n ? j : n - 1;
                    const double bb = BM == 2 ? beta[jc] : beta_scalar;
                    const double y = Y[rowc + (int64_t)jc * ldy];
                    init[v] = (rok[a] && j < n && bb != 0.0) ? bb * y : 0.0;
                }
            }
            // D = 0 * 0 + init moves the value into the accumulator registers
            acc[a][q] = BM == 0 ? init : mfma(0.0, 0.0, init);
        }
        if (BM != 0) asm volatile("" ::: "memory");
    }
    // bx[b][a]: X operand of row tile a, ac[b][q]: C^T operand of column
    // tile q, for the step held in ring buffer b; steps past m load zeros
    double bx[kRing][4], ac[kRing][4];
    // one step into ring buffer B (a macro: the ring index must stay a
    // compile-time constant for the buffers to live in registers)
#define GCGE_LC_LOAD(B, STEP)                                                         \
    {                                                                                 \
        const int k_ = (STEP) * 4 + lk;                                               \
        const bool kok_ = k_ < m;                                                     \
        _Pragma("unroll") for (int a = 0; a < 4; ++a) bx[B][a] =                      \
            (kok_ && rok[a]) ? xr[a][(int64_t)k_ * ldx] : 0.0;                        \
        const double *cr_ = Ct + (int64_t)(kok_ ? k_ : 0) * ldct + j0 + li;           \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) ac[B][q] =                      \
            (kok_ && jok[q]) ? cr_[q * 16] : 0.0;                                     \
    }
#pragma unroll
    for (int b = 0; b < kRing - 1; ++b) GCGE_LC_LOAD(b, b);
    const int nsteps = (m + 3) / 4;
    for (int s0 = 0; s0 < nsteps; s0 += kRing) {
#pragma unroll
        for (int u = 0; u < kRing; ++u) {
            // refill the buffer the previous step used
            GCGE_LC_LOAD((u + kRing - 1) % kRing, s0 + u + kRing - 1);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[a][q] = mfma(ac[u][q], bx[u][a], acc[a][q]);
        }
    }
#undef GCGE_LC_LOAD
    // lane l holds (Y^T)[j0 + q*16 + l/16 + 4v][i0 + a*16 + l%16]
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (!rok[a]) continue;
        const int64_t row = i0 + a * 16 + li;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + q * 16 + lk + 4 * v;
                if (j < n) Y[row + (int64_t)j * ldy] = acc[a][q][v];
            }
    }
}

}  // namespace

int gram_mfma_choice(int64_t nrows, int m, int n, bool lower_only, bool *tune)
{
    const int ch = tune_lookup(gram_key(nrows, m, n, lower_only));
    if (tune) *tune = ch < 0 && tune_enabled();
    return (ch < 0 || ch >= kGramCands) ? kGramDefault : ch;
}

// choice -1: at most max_chunks row chunks of at least kGramMinRows[0] rows
static GramPlanM plan_of(int64_t nrows, int m, int n, int choice, int64_t max_chunks)
{
    if (choice >= 0) return plan_cand(nrows, m, n, choice);
    const int64_t tiles = (int64_t)((m + kGT - 1) / kGT) * ((n + kGT - 1) / kGT);
    return plan_with(nrows, m, n, kGramMinRows[0], (max_chunks < 1 ? 1 : max_chunks) * tiles);
}

int64_t gram_mfma_chunk_doubles(int m, int n)
{
    return (int64_t)((m + kGT - 1) / kGT) * ((n + kGT - 1) / kGT) * kGT * kGT;
}

int64_t gram_mfma_scratch(int64_t nrows, int m, int n, int choice, int64_t max_chunks)
{
    const GramPlanM p = plan_of(nrows, m, n, choice, max_chunks);
    return (int64_t)p.nchunk * p.tm * p.tn * kGT * kGT;
}

namespace {

void launch_gram(const GramPlanM &p, int64_t nrows, int m, int n, const double *X, int64_t ldx,
                 const double *Y, int64_t ldy, double *G, int ldg, bool lower_only,
                 double *partial, hipStream_t s)
{
    const int64_t nblk = (int64_t)p.nchunk * p.tm * p.tn;
    k_gram_mfma<<<(unsigned)nblk, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.tm, p.tn,
                                               p.rows_per_chunk, lower_only ? 1 : 0, partial);
    const int64_t total = (int64_t)p.tm * p.tn * kGT * kGT;
    k_gram_mfma_reduce<<<(unsigned)((total + kWG - 1) / kWG), kWG, 0, s>>>(
        m, n, p.tm, p.tn, p.nchunk, partial, G, ldg);
}

}  // namespace

void gram_mfma(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Y,
               int64_t ldy, double *G, int ldg, bool lower_only, int choice, int64_t max_chunks,
               bool tune, double *partial, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    if (tune) {
        // time every chunking on a private partial buffer large enough for
        // all of them (the caller sized its buffer for the default plan)
        const TuneKey key = gram_key(nrows, m, n, lower_only);
        int64_t need = 0;
        for (int i = 0; i < kGramCands; ++i) {
            const GramPlanM q = plan_cand(nrows, m, n, i);
            const int64_t e = (int64_t)q.nchunk * q.tm * q.tn * kGT * kGT;
            if (e > need) need = e;
        }
        double *buf = nullptr;
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&buf), sizeof(double) * (size_t)need));
        tune_run(key, kGramCands, kGramDefault,
                 [&](int i) {
                     launch_gram(plan_cand(nrows, m, n, i), nrows, m, n, X, ldx, Y, ldy, G, ldg,
                                 lower_only, buf, s);
                 },
                 s);
        HIPCHECK(hipStreamSynchronize(s));
        HIPCHECK(hipFree(buf));
        return;
    }
    launch_gram(plan_of(nrows, m, n, choice, max_chunks), nrows, m, n, X, ldx, Y, ldy, G, ldg,
                lower_only, partial, s);
}

void lincomb_mfma(int64_t nrows, int m, int n, const double *X, int64_t ldx, const double *Ct,
                  int ldct, int beta_mode, double beta_scalar, const double *beta, double *Y,
                  int64_t ldy, hipStream_t s)
{
    if (nrows <= 0 || n <= 0) return;
    const int ntj = (n + 63) / 64;
    const int64_t nrt = (nrows + 255) / 256;
    const unsigned g = (unsigned)(nrt * ntj);
    const bool deep = nrows >= kRingDeepRows;
#define GCGE_LC_LAUNCH(BM)                                                                  \
    do {                                                                                    \
        if (deep)                                                                           \
            k_lincomb_mfma<BM, 5><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Ct, ldct,          \
                                                    beta_scalar, beta, Y, ldy, ntj);        \
        else                                                                                \
            k_lincomb_mfma<BM, 4><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Ct, ldct,          \
                                                    beta_scalar, beta, Y, ldy, ntj);        \
    } while (0)
    if (beta_mode == 0)
        GCGE_LC_LAUNCH(0);
    else if (beta_mode == 1)
        GCGE_LC_LAUNCH(1);
    else
        GCGE_LC_LAUNCH(2);
#undef GCGE_LC_LAUNCH
}

}  // namespace gcge

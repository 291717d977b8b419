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
// Gram: G (m x n) = X^T Y.  A workgroup owns a TM x TN tile of G over one
// chunk of rows; its four waves, WM along m and 4 / WM along n, own 64 x
// 16 NB blocks (4 x NB MFMA tiles).  The default shape is 128 x 128 (WM 2,
// NB 4); a Y side of at most 64 columns -- the block width of most
// orthogonalization Grams -- takes 256 x 16 NB tiles (WM 4, NB = its
// 16-column tiles), and both sides of at most 64 a 64 x 64 tile (WM 1, NB
// 1), so the matrix cores do not multiply 128-wide zero padding.
// Rows are staged 16 at a time through LDS, stored column-major (row index
// fastest, padded to 17) so each 16-lane MFMA operand read covers the
// banks once.  Each thread fetches 4 consecutive rows (one 32-byte load)
// of its columns of each operand, and the next stage's loads are issued
// into registers before the current stage's MFMAs, so HBM latency overlaps
// the matrix-core work; the default shape holds 98 VGPRs + 128 AGPRs, two
// waves per SIMD.  Partial tiles per chunk are summed by
// k_gram_mfma_reduce in a fixed order (deterministic, no atomics).  Each
// entry sums its rows in the same order whatever the shape; the chunking,
// and so the rounding of the partial sums, follows the tile count.
// ---------------------------------------------------------------------
constexpr int kGK = 16;        // rows per stage
constexpr int kGKP = kGK + 1;
constexpr int kGRG = kGK / 4;  // 4-row groups per stage

template <int WM, int NB>
__global__ __launch_bounds__(kWG) void k_gram_mfma(
    int64_t nrows, int m, int n, const double *__restrict__ X, int64_t ldx,
    const double *__restrict__ Y, int64_t ldy, int tm, int tn, int64_t rows_per_chunk,
    int lower_only, double *__restrict__ partial)
{
    constexpr int WN = 4 / WM;
    constexpr int TM = WM * 64, TN = WN * 16 * NB;
    constexpr int XC = TM * kGRG / kWG;                  // X columns per thread
    constexpr int YC = (TN * kGRG + kWG - 1) / kWG;      // Y columns per thread
    __shared__ double sx[TM * kGKP];
    __shared__ double sy[TN * kGKP];
    const int64_t lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tiles = tm * tn;
    const int chunk = (int)(lb / tiles);
    const int t = (int)(lb % tiles);
    const int bi = t / tn, bj = t % tn;
    double *out = partial + ((int64_t)chunk * tiles + t) * (TM * TN);
    const int i0 = bi * TM, j0 = bj * TN;
    if (lower_only && i0 + TM - 1 < j0) {  // wholly above the diagonal
        for (int e = threadIdx.x; e < TM * TN; e += kWG) out[e] = 0.0;
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t rbeg = (int64_t)chunk * rows_per_chunk;
    int64_t rend = rbeg + rows_per_chunk;
    if (rend > nrows) rend = nrows;
    // this thread's slots in a stage: rows 4*rg .. 4*rg+3 of its columns
    // (one 32-byte load per column; kGRG lanes cover the stage's rows of a
    // column, so a wave instruction reads whole row segments of 64/kGRG
    // columns, and only the thread's column addresses stay live)
    const int rg = threadIdx.x % kGRG;
    const int cg = threadIdx.x / kGRG;
    const double *xc[XC], *yc[YC];
    bool xok[XC], yok[YC];
#pragma unroll
    for (int c = 0; c < XC; ++c) {
        const int cc = cg * XC + c;
        xok[c] = i0 + cc < m;
        xc[c] = X + (int64_t)(xok[c] ? i0 + cc : 0) * ldx + 4 * rg;
    }
#pragma unroll
    for (int c = 0; c < YC; ++c) {
        const int cc = cg * YC + c;
        yok[c] = cc < TN && j0 + cc < n;
        yc[c] = Y + (int64_t)(yok[c] ? j0 + cc : 0) * ldy + 4 * rg;
    }
    double px[XC][4], py[YC][4];
    auto fetch = [&](int64_t r0) {
        const int64_t r = r0 + 4 * rg;
        if (r + 4 <= rend) {
#pragma unroll
            for (int c = 0; c < XC; ++c) {
                const d4 v = xok[c] ? *reinterpret_cast<const d4 *>(xc[c] + r0)
                                    : d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) px[c][q] = v[q];
            }
#pragma unroll
            for (int c = 0; c < YC; ++c) {
                const d4 v = yok[c] ? *reinterpret_cast<const d4 *>(yc[c] + r0)
                                    : d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) py[c][q] = v[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool rok = r + q < rend;
#pragma unroll
                for (int c = 0; c < XC; ++c) px[c][q] = (rok && xok[c]) ? xc[c][r0 + q] : 0.0;
#pragma unroll
                for (int c = 0; c < YC; ++c) py[c][q] = (rok && yok[c]) ? yc[c][r0 + q] : 0.0;
            }
        }
    };
    d4 acc[4][NB];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    if (rbeg < rend) fetch(rbeg);
    for (int64_t r0 = rbeg; r0 < rend; r0 += kGK) {
#pragma unroll
        for (int c = 0; c < XC; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) sx[(cg * XC + c) * kGKP + 4 * rg + q] = px[c][q];
#pragma unroll
        for (int c = 0; c < YC; ++c) {
            if (cg * YC + c < TN) {
#pragma unroll
                for (int q = 0; q < 4; ++q) sy[(cg * YC + c) * kGKP + 4 * rg + q] = py[c][q];
            }
        }
        __syncthreads();
        if (r0 + kGK < rend) fetch(r0 + kGK);
#pragma unroll 2
        for (int kk = 0; kk < kGK / 4; ++kk) {
            const int row = kk * 4 + lk;
            double av[4], bv[NB];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sx[(wm * 64 + a * 16 + li) * kGKP + row];
#pragma unroll
            for (int b = 0; b < NB; ++b) bv[b] = sy[(wn * 16 * NB + b * 16 + li) * kGKP + row];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[a][b] = mfma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = wm * 64 + a * 16 + lk + 4 * v;
                const int jj = wn * 16 * NB + b * 16 + li;
                out[ii * TN + jj] = acc[a][b][v];
            }
}

__global__ __launch_bounds__(kWG) void k_gram_mfma_reduce(int m, int n, int tm, int tn, int TM,
                                                          int TN, int nchunk,
                                                          const double *__restrict__ partial,
                                                          double *__restrict__ G, int ldg)
{
    // threads walk the partial tiles in storage order, so each chunk's read
    // is contiguous across the wavefront
    const int64_t off = (int64_t)blockIdx.x * kWG + threadIdx.x;
    const int64_t stride = (int64_t)tm * tn * TM * TN;
    if (off >= stride) return;
    const int t = (int)(off / (TM * TN));
    const int within = (int)(off % (TM * TN));
    const int i = (t / tn) * TM + within / TN, j = (t % tn) * TN + within % TN;
    if (i >= m || j >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[off + (int64_t)c * stride];
    G[i + (int64_t)j * ldg] = s;
}

// Tile shape of a Gram of m x n (see k_gram_mfma).
struct GramShape {
    int wm, nb;
    int rows() const { return wm * 64; }
    int cols() const { return (4 / wm) * 16 * nb; }
};

GramShape gram_shape(int m, int n)
{
    if (m <= 64 && n <= 64) return GramShape{1, 1};
    if (n <= 64) return GramShape{4, (n + 15) / 16};
    return GramShape{2, 4};
}

struct GramPlanM {
    GramShape sh;
    int tm, tn, nchunk;
    int64_t rows_per_chunk;
    int64_t tile_doubles() const { return (int64_t)sh.rows() * sh.cols(); }
};

int64_t gram_tiles(int m, int n)
{
    const GramShape sh = gram_shape(m, n);
    return (int64_t)((m + sh.rows() - 1) / sh.rows()) * ((n + sh.cols() - 1) / sh.cols());
}

// Row chunking: at most max_wg workgroups in total (256 CUs, two resident
// each, at the default 512: one more than that runs alone in a second
// round and doubles the time) and at least min_rows rows (min_rows / 16
// stages) per chunk, so small problems still spread over many workgroups
// instead of a few walking all their rows.
GramPlanM plan_with(int64_t nrows, int m, int n, int64_t min_rows, int64_t max_wg)
{
    GramPlanM p;
    p.sh = gram_shape(m, n);
    p.tm = (m + p.sh.rows() - 1) / p.sh.rows();
    p.tn = (n + p.sh.cols() - 1) / p.sh.cols();
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
    const int64_t tiles = gram_tiles(m, n);
    return TuneKey{kTuneGram, ilog2_floor(nrows), tiles < 64 ? (int)tiles : 64,
                   lower_only ? 1 : 0};
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

// NB: 16-column MFMA tiles per wave (4, a 64-column tile, unless the
// block is narrower: 40 destination columns run 48 wide instead of 64).
template <int BM, int kRing, int NB>
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
    const int j0 = tj * 16 * NB;
    bool rok[4];
    const double *xr[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t row = i0 + a * 16 + li;
        rok[a] = row < nrows;
        xr[a] = X + (rok[a] ? row : 0);
    }
    bool jok[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) jok[q] = (j0 + q * 16 + li) < n;
    // BM (beta mode) 0: overwrite; 1: scalar beta; 2: per-column beta.  The
    // accumulators start from Y diag(beta), read 16 at a time (a scheduling
    // barrier after each row tile keeps the 64 loads from being in flight at
    // once, which would double the registers and halve the occupancy)
    d4 acc[4][NB];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        // out-of-range rows and columns read a clamped in-range element and
        // are zeroed by the select (no branch per element)
        const int64_t rowc = rok[a] ? i0 + a * 16 + li : nrows - 1;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
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
    double bx[kRing][4], ac[kRing][NB];
    // one step into ring buffer B (a macro: the ring index must stay a
    // compile-time constant for the buffers to live in registers)
#define GCGE_LC_LOAD(B, STEP)                                                         \
    {                                                                                 \
        const int k_ = (STEP) * 4 + lk;                                               \
        const bool kok_ = k_ < m;                                                     \
        _Pragma("unroll") for (int a = 0; a < 4; ++a) bx[B][a] =                      \
            (kok_ && rok[a]) ? xr[a][(int64_t)k_ * ldx] : 0.0;                        \
        const double *cr_ = Ct + (int64_t)(kok_ ? k_ : 0) * ldct + j0 + li;           \
        _Pragma("unroll") for (int q = 0; q < NB; ++q) ac[B][q] =                     \
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
                for (int q = 0; q < NB; ++q) acc[a][q] = mfma(ac[u][q], bx[u][a], acc[a][q]);
        }
    }
#undef GCGE_LC_LOAD
    // lane l holds (Y^T)[j0 + q*16 + l/16 + 4v][i0 + a*16 + l%16]
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (!rok[a]) continue;
        const int64_t row = i0 + a * 16 + li;
#pragma unroll
        for (int q = 0; q < NB; ++q)
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
    return plan_with(nrows, m, n, kGramMinRows[0],
                     (max_chunks < 1 ? 1 : max_chunks) * gram_tiles(m, n));
}

int64_t gram_mfma_chunk_doubles(int m, int n)
{
    const GramShape sh = gram_shape(m, n);
    return gram_tiles(m, n) * sh.rows() * sh.cols();
}

int64_t gram_mfma_scratch(int64_t nrows, int m, int n, int choice, int64_t max_chunks)
{
    const GramPlanM p = plan_of(nrows, m, n, choice, max_chunks);
    return (int64_t)p.nchunk * p.tm * p.tn * p.tile_doubles();
}

namespace {

void launch_gram(const GramPlanM &p, int64_t nrows, int m, int n, const double *X, int64_t ldx,
                 const double *Y, int64_t ldy, double *G, int ldg, bool lower_only,
                 double *partial, hipStream_t s)
{
    const int64_t nblk = (int64_t)p.nchunk * p.tm * p.tn;
#define GCGE_GRAM(WM, NB)                                                                      \
    k_gram_mfma<WM, NB><<<(unsigned)nblk, kWG, 0, s>>>(nrows, m, n, X, ldx, Y, ldy, p.tm, p.tn, \
                                                       p.rows_per_chunk, lower_only ? 1 : 0,   \
                                                       partial)
    if (p.sh.wm == 1)
        GCGE_GRAM(1, 1);
    else if (p.sh.wm == 4 && p.sh.nb == 1)
        GCGE_GRAM(4, 1);
    else if (p.sh.wm == 4 && p.sh.nb == 2)
        GCGE_GRAM(4, 2);
    else if (p.sh.wm == 4 && p.sh.nb == 3)
        GCGE_GRAM(4, 3);
    else if (p.sh.wm == 4)
        GCGE_GRAM(4, 4);
    else
        GCGE_GRAM(2, 4);
#undef GCGE_GRAM
    const int64_t total = (int64_t)p.tm * p.tn * p.tile_doubles();
    k_gram_mfma_reduce<<<(unsigned)((total + kWG - 1) / kWG), kWG, 0, s>>>(
        m, n, p.tm, p.tn, p.sh.rows(), p.sh.cols(), p.nchunk, partial, G, ldg);
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
            const int64_t e = (int64_t)q.nchunk * q.tm * q.tn * q.tile_doubles();
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
    const int nb = n <= 16 ? 1 : n <= 32 ? 2 : n <= 48 ? 3 : 4;
    const int ntj = (n + 16 * nb - 1) / (16 * nb);
    const int64_t nrt = (nrows + 255) / 256;
    const unsigned g = (unsigned)(nrt * ntj);
    const bool deep = nrows >= kRingDeepRows;
#define GCGE_LC_LAUNCH(BM, NB)                                                              \
    do {                                                                                    \
        if (deep)                                                                           \
            k_lincomb_mfma<BM, 5, NB><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Ct, ldct,      \
                                                        beta_scalar, beta, Y, ldy, ntj);    \
        else                                                                                \
            k_lincomb_mfma<BM, 4, NB><<<g, kWG, 0, s>>>(nrows, m, n, X, ldx, Ct, ldct,      \
                                                        beta_scalar, beta, Y, ldy, ntj);    \
    } while (0)
#define GCGE_LC_NB(BM)                                                                      \
    do {                                                                                    \
        if (nb == 1)                                                                        \
            GCGE_LC_LAUNCH(BM, 1);                                                          \
        else if (nb == 2)                                                                   \
            GCGE_LC_LAUNCH(BM, 2);                                                          \
        else if (nb == 3)                                                                   \
            GCGE_LC_LAUNCH(BM, 3);                                                          \
        else                                                                                \
            GCGE_LC_LAUNCH(BM, 4);                                                          \
    } while (0)
    if (beta_mode == 0)
        GCGE_LC_NB(0);
    else if (beta_mode == 1)
        GCGE_LC_NB(1);
    else
        GCGE_LC_NB(2);
#undef GCGE_LC_NB
#undef GCGE_LC_LAUNCH
}

}  // namespace gcge

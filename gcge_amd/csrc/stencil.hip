// Matrix-free products with constant-coefficient 7-point grid operators.
//
// The 3D Laplacian and mass matrices of the stencil problems (A = 6 I - Adj,
// B = (12 I + Adj) / 18 on an n^3 Dirichlet grid, point x + n y + n^2 z) have
// one diagonal value d and one neighbour value o.  For them MatDotMultiVec
// (app/app_ccs.c:50-139) needs no matrix reads and no gathers: a wavefront
// holds one x-line of the grid (lane l: points x = 2l and 2l + 1) for kTY
// consecutive y-lines, and marches along z with the planes z - 1, z, z + 1
// of its points in registers.  Per plane step it loads plane z + 1 of its
// lines and the two lines just outside them (y0 - 1 and y0 + kTY, the
// points of the wavefronts next to it), takes x +- 1 from its own lane or
// the adjacent ones (cross-lane moves), and writes its kTY output lines: a
// column's elements are loaded once plus the y and z halos, against seven
// gathers per element in the SELL row kernel.
//
// Each output is summed in the order of the SELL row (ascending column:
// -z, -y, -x, centre, +x, +y, +z, the in-grid ones only) with one fused
// multiply-add per entry from 0, which is what the SELL kernel computes for
// the same matrix, so the results are the same bits (padding entries of the
// SELL layout add 0 * x, which leaves every nonzero sum unchanged).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "app_hip_internal.h"
#include "kernels.h"
#include "sell_rows.h"

namespace gcge {
namespace {

constexpr int kZChunk = 32;  // planes marched by one workgroup

__device__ __forceinline__ double lane_up(double v)
{
    // the value of lane l - 1 (lane 0 gets its own; callers do not use it)
    return __shfl_up(v, 1, 64);
}

__device__ __forceinline__ double lane_down(double v)
{
    return __shfl_down(v, 1, 64);
}

// Y[:, c] = (d I + o Adj) X[:, c] on the n^3 grid, columns c < ncols.
// Workgroup: 4 wavefronts, wavefront w holds y-lines (4 g + w) kTY ...
// of z-planes [q kZChunk, (q + 1) kZChunk) of column c.  VEC: n even, so
// a lane's two points are one aligned 16-byte access.
template <int kTY, bool VEC>
__global__ __launch_bounds__(256) void k_stencil7(int n, int ygroups, int zchunks, double d,
                                                  double o, const double *__restrict__ X,
                                                  int64_t ldx, double *__restrict__ Y,
                                                  int64_t ldy)
{
    const int64_t L = row_block<true>();  // XCD-contiguous: neighbouring tiles share an L2
    const int g = (int)(L % ygroups);
    const int64_t t = L / ygroups;
    const int q = (int)(t % zchunks);
    const int64_t c = t / zchunks;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int y0 = (g * 4 + wv) * kTY;
    if (y0 >= n) return;  // wave-uniform
    const int x0 = 2 * lane;
    const bool a0 = x0 < n, a1 = x0 + 1 < n;
    const int z0 = q * kZChunk;
    const int z1 = z0 + kZChunk < n ? z0 + kZChunk : n;
    const int64_t nn = (int64_t)n * n;
    const double *Xc = X + c * ldx;
    double *Yc = Y + c * ldy;

    // points x0, x0 + 1 of line (y, z); zeros outside the grid
    auto load2 = [&](int y, int z, double &u, double &v) {
        u = 0.0;
        v = 0.0;
        if (y < 0 || y >= n || z < 0 || z >= n) return;  // wave-uniform
        const int64_t r = x0 + (int64_t)n * y + nn * z;
        if constexpr (VEC) {
            if (a0) {
                const double2 p = *reinterpret_cast<const double2 *>(Xc + r);
                u = p.x;
                v = p.y;
            }
        } else {
            if (a0) u = Xc[r];
            if (a1) v = Xc[r + 1];
        }
    };

    double pm[kTY][2], pc[kTY][2], pn[kTY][2];
#pragma unroll
    for (int i = 0; i < kTY; ++i) {
        load2(y0 + i, z0 - 1, pm[i][0], pm[i][1]);
        load2(y0 + i, z0, pc[i][0], pc[i][1]);
    }
    for (int z = z0; z < z1; ++z) {
        double hm0, hm1, hp0, hp1;
#pragma unroll
        for (int i = 0; i < kTY; ++i) load2(y0 + i, z + 1, pn[i][0], pn[i][1]);
        load2(y0 - 1, z, hm0, hm1);
        load2(y0 + kTY, z, hp0, hp1);
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            const int y = y0 + i;
            if (y >= n) break;  // wave-uniform
            // x - 1 of point 0 is point 1 of the lane before; x + 1 of
            // point 1 is point 0 of the lane after
            const double left = lane_up(pc[i][1]);
            const double right = lane_down(pc[i][0]);
            double out[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int x = x0 + p;
                const double xm = p == 0 ? left : pc[i][0];
                const double xp = p == 0 ? pc[i][1] : right;
                const double ym = i > 0 ? pc[i - 1][p] : (p == 0 ? hm0 : hm1);
                const double yp = i < kTY - 1 ? pc[i + 1][p] : (p == 0 ? hp0 : hp1);
                double acc = 0.0;
                if (z > 0) acc = __builtin_fma(o, pm[i][p], acc);
                if (y > 0) acc = __builtin_fma(o, ym, acc);
                if (x > 0) acc = __builtin_fma(o, xm, acc);
                acc = __builtin_fma(d, pc[i][p], acc);
                if (x < n - 1) acc = __builtin_fma(o, xp, acc);
                if (y < n - 1) acc = __builtin_fma(o, yp, acc);
                if (z < n - 1) acc = __builtin_fma(o, pn[i][p], acc);
                out[p] = acc;
            }
            const int64_t r = x0 + (int64_t)n * y + nn * z;
            if constexpr (VEC) {
                if (a0) *reinterpret_cast<double2 *>(Yc + r) = make_double2(out[0], out[1]);
            } else {
                if (a0) Yc[r] = out[0];
                if (a1) Yc[r + 1] = out[1];
            }
        }
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            pm[i][0] = pc[i][0];
            pm[i][1] = pc[i][1];
            pc[i][0] = pn[i][0];
            pc[i][1] = pn[i][1];
        }
    }
}

template <int kTY>
void launch_stencil7(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
                     int ncols, hipStream_t s)
{
    const int n = A.stencil_n;
    const int ygroups = (n + 4 * kTY - 1) / (4 * kTY);
    const int zchunks = (n + kZChunk - 1) / kZChunk;
    const int64_t blocks = (int64_t)ygroups * zchunks * ncols;
    const bool vec = (n % 2 == 0) && (ldx % 2 == 0) && (ldy % 2 == 0) &&
                     (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
    if (vec)
        k_stencil7<kTY, true><<<(unsigned)blocks, 256, 0, s>>>(n, ygroups, zchunks, A.stencil_d,
                                                               A.stencil_o, X, ldx, Y, ldy);
    else
        k_stencil7<kTY, false><<<(unsigned)blocks, 256, 0, s>>>(n, ygroups, zchunks, A.stencil_d,
                                                                A.stencil_o, X, ldx, Y, ldy);
}

// ---------------------------------------------------------------------
// The FP32 BlockPCG product w = A p over its 8-column tiles (element
// (row, c) of tile t at ((t n + row) 8 + c % 8), one 32-byte row), with the
// partial p.w sums, for a constant-coefficient 7-point A without shift.
// A lane holds one grid point (all 8 columns of its row); ceil(n / 64)
// wavefronts hold an x-line, the other wavefronts of the workgroup further
// y-lines; each marches kTY lines along z with planes z - 1, z, z + 1 in
// registers.  The x +- 1 rows are loaded directly (the same cache lines as
// the centre row, shifted by one row); the far neighbours (y +- 1, z +- 1)
// come from registers, except the two halo lines.  Each output is summed
// in the SELL row's order with one fused multiply-add per entry.  Partial
// p.w: each thread accumulates over its points, the workgroup sums its
// threads in a fixed tree, and partial[(tile 8 + t) nparts + part] holds
// workgroup `part`'s sum.
// ---------------------------------------------------------------------
struct alignas(32) f8s {
    float v[8];
};

__device__ __forceinline__ f8s load_row8(const float *p)
{
    return *reinterpret_cast<const f8s *>(p);
}

template <int kTY>
__global__ __launch_bounds__(256) void k_stencil_pw(int n, int xw, int ygroups, int zchunks,
                                                    float d, float o,
                                                    const int *__restrict__ tiles,
                                                    const int *__restrict__ ntact_p,
                                                    const float *__restrict__ P,
                                                    float *__restrict__ Wv,
                                                    float *__restrict__ partial, int nparts)
{
    __shared__ float red[(256 / 64 + 1) * 8];
    const int k = blockIdx.y;
    if (k >= *ntact_p) return;  // workgroup-uniform
    const int tile = tiles[k];
    const int64_t L = row_block<true>();
    const int g = (int)(L % ygroups);
    const int q = (int)(L / ygroups);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int segs = 4 / xw;
    const int y0 = (g * segs + wv / xw) * kTY;
    const int x = (wv % xw) * 64 + lane;
    const bool livex = x < n;
    const bool wave_on = y0 < n;  // wave-uniform
    const int z0 = q * kZChunk;
    const int z1 = z0 + kZChunk < n ? z0 + kZChunk : n;
    const int64_t nn = (int64_t)n * n;
    const int64_t N = nn * n;
    const float *Pt = P + (int64_t)tile * N * 8;
    float *Wt = Wv + (int64_t)tile * N * 8;

    // row (x + dx, y, z) of the tile; zeros outside the grid
    auto load = [&](int dx, int y, int z) {
        f8s r;
#pragma unroll
        for (int t = 0; t < 8; ++t) r.v[t] = 0.f;
        const int xx = x + dx;
        if (y < 0 || y >= n || z < 0 || z >= n) return r;  // wave-uniform
        if (livex && xx >= 0 && xx < n) r = load_row8(Pt + (xx + (int64_t)n * y + nn * z) * 8);
        return r;
    };

    float dot[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) dot[t] = 0.f;
    if (wave_on) {
        f8s pm[kTY], pc[kTY], pn[kTY];
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            pm[i] = load(0, y0 + i, z0 - 1);
            pc[i] = load(0, y0 + i, z0);
        }
        for (int z = z0; z < z1; ++z) {
#pragma unroll
            for (int i = 0; i < kTY; ++i) pn[i] = load(0, y0 + i, z + 1);
            const f8s hm = load(0, y0 - 1, z), hp = load(0, y0 + kTY, z);
#pragma unroll
            for (int i = 0; i < kTY; ++i) {
                const int y = y0 + i;
                if (y >= n) break;  // wave-uniform
                const f8s left = load(-1, y, z), right = load(1, y, z);
                const f8s &ym = i > 0 ? pc[i - 1] : hm;
                const f8s &yp = i < kTY - 1 ? pc[i + 1] : hp;
                f8s w;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    float acc = 0.f;
                    if (z > 0) acc = __builtin_fmaf(o, pm[i].v[t], acc);
                    if (y > 0) acc = __builtin_fmaf(o, ym.v[t], acc);
                    if (x > 0) acc = __builtin_fmaf(o, left.v[t], acc);
                    acc = __builtin_fmaf(d, pc[i].v[t], acc);
                    if (x < n - 1) acc = __builtin_fmaf(o, right.v[t], acc);
                    if (y < n - 1) acc = __builtin_fmaf(o, yp.v[t], acc);
                    if (z < n - 1) acc = __builtin_fmaf(o, pn[i].v[t], acc);
                    w.v[t] = acc;
                    dot[t] += acc * pc[i].v[t];
                }
                if (livex)
                    *reinterpret_cast<f8s *>(Wt + (x + (int64_t)n * y + nn * z) * 8) = w;
            }
#pragma unroll
            for (int i = 0; i < kTY; ++i) {
                pm[i] = pc[i];
                pc[i] = pn[i];
            }
        }
    }
    // the workgroup's sum of its threads' partial dots, in a fixed order
    const int ln = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float v = dot[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (ln == 0) red[wv * 8 + t] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < 4; ++w) s += red[w * 8 + threadIdx.x];
        partial[(int64_t)(tile * 8 + threadIdx.x) * nparts + L] = s;
    }
}

constexpr int kPwTY = 4;

}  // namespace

int stencil_pw_parts(const SellMat &A)
{
    const int n = A.stencil_n;
    if (n < 4 || n > kStencilMaxN) return 0;
    const int xw = (n + 63) / 64;
    const int ygroups = (n + (4 / xw) * kPwTY - 1) / ((4 / xw) * kPwTY);
    return ygroups * ((n + kZChunk - 1) / kZChunk);
}

void stencil_pw(const SellMat &A, const int *tiles, const int *ntact_p, int nt_bound,
                const float *P, float *W, float *partial, hipStream_t s)
{
    const int n = A.stencil_n;
    const int xw = (n + 63) / 64;
    const int ygroups = (n + (4 / xw) * kPwTY - 1) / ((4 / xw) * kPwTY);
    const int zchunks = (n + kZChunk - 1) / kZChunk;
    const int parts = ygroups * zchunks;
    k_stencil_pw<kPwTY><<<dim3((unsigned)parts, (unsigned)nt_bound), 256, 0, s>>>(
        n, xw, ygroups, zchunks, (float)A.stencil_d, (float)A.stencil_o, tiles, ntact_p, P, W,
        partial, parts);
}

bool spmm_stencil(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
                  int ncols, hipStream_t s)
{
    if (A.stencil_n < 4 || A.stencil_n > kStencilMaxN || ncols <= 0) return false;
    launch_stencil7<8>(A, X, ldx, Y, ldy, ncols, s);
    return true;
}

}  // namespace gcge

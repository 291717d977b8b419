// Matrix-free products with constant-coefficient 7-point grid operators.
//
// The 3D Laplacian and mass matrices of the stencil problems (A = 6 I - Adj,
// B = (12 I + Adj) / 18 on an n^3 Dirichlet grid, point x + n y + n^2 z) have
// one diagonal value d and one neighbour value o.  For them MatDotMultiVec
// (app/app_ccs.c:50-139) needs no matrix reads and no gathers: a wavefront
// holds one x-line of the grid (lane l: points x = 2l and 2l + 1) for kTY
// consecutive y-lines, and marches along z with the planes z - 1, z, z + 1
// of its points in registers and plane z + 2 on its way.  It takes x +- 1
// from its own lane or the adjacent ones (cross-lane moves), y +- 1 from
// its registers or, across wavefronts, from LDS, and writes its kTY output
// lines: a column's elements are loaded once plus the workgroup's y and z
// halos (mostly L2 hits), against seven gathers per element in the SELL
// row kernel.  Short marches over many workgroups measured fastest: the
// loads in flight, not the halo re-reads, set the rate.
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
// Workgroup: kW wavefronts, wavefront w holds y-lines (kW g + w) kTY ...
// of z-planes [q zc, (q + 1) zc) of column c.  VEC: n even, so a lane's
// two points are one aligned 16-byte access.  The y-halo lines of a
// wavefront are the first and last lines of the wavefronts beside it in
// the same workgroup, exchanged through LDS each plane step (two slots, by
// plane parity, so one barrier per step suffices); only the workgroup's
// two outer halo lines are loaded from memory.  Loads run one plane ahead:
// a wavefront issues the loads of plane z + 2 (and of its outer halo line
// of plane z + 1) before it computes plane z, so they are in flight during
// the computation and the exchange; four plane register sets rotate by
// unrolling the march by four, so no register move waits on a load.
// RESID: Y[:, c] = A X[:, c] - lam[c] (B X[:, c]) for a second operator B =
// dB I + oB Adj on the same grid, each product summed as above, then the
// scaling and the difference rounded separately -- the bits of the product
// by A, the product by B scaled by lam[c] (0 when lam[c] is 0) and the
// difference, as three separate kernels give them.
__device__ __forceinline__ double resid_of(double ax, double bx, double lam)
{
#pragma clang fp contract(off)
    const double t = lam == 0.0 ? 0.0 : lam * bx;
    return ax - t;
}

template <int kTY, int kW, bool VEC, bool RESID>
__global__ __launch_bounds__(64 * kW) void k_stencil7(int n, int ygroups, int zchunks, int zc,
                                                   double d, double o, double dB, double oB,
                                                   const double *__restrict__ lam,
                                                   const double *__restrict__ X, int64_t ldx,
                                                   double *__restrict__ Y, int64_t ldy)
{
    __shared__ double2 s_edge[2][2][kW][64];  // [parity][first, last][wavefront][lane]
    const int64_t L = row_block<true>();  // XCD-contiguous: neighbouring tiles share an L2
    const int g = (int)(L % ygroups);
    const int64_t t = L / ygroups;
    const int q = (int)(t % zchunks);
    const int64_t c = t / zchunks;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int y0 = (g * kW + wv) * kTY;
    // waves past the grid's last line still meet the barriers
    const bool wave_on = y0 < n;  // wave-uniform
    const int x0 = 2 * lane;
    const bool a0 = x0 < n, a1 = x0 + 1 < n;
    const int z0 = q * zc;
    const int z1 = z0 + zc < n ? z0 + zc : n;
    const int64_t nn = (int64_t)n * n;
    const double *Xc = X + c * ldx;
    double *Yc = Y + c * ldy;
    const double lam_c = RESID ? lam[c] : 0.0;
    // the outer halo line of this wavefront (wavefronts 0 and 3 only)
    const int yh = wv == 0 ? y0 - 1 : y0 + kTY;
    const bool outer = wave_on && (wv == 0 || wv == kW - 1);

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
    using Plane = double[kTY][2];
    auto load_plane = [&](Plane &p, int z) {
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            p[i][0] = p[i][1] = 0.0;
            if (wave_on) load2(y0 + i, z, p[i][0], p[i][1]);
        }
    };
    auto load_halo = [&](double (&h)[2], int z) {
        h[0] = h[1] = 0.0;
        if (outer) load2(yh, z, h[0], h[1]);
    };
    // plane z from pm (z - 1), pc (z), pn (z + 1) and the outer halo hc of
    // plane z; loads plane z + 2 into pf and the halo of z + 1 into hn
    auto step = [&](int z, const Plane &pm, const Plane &pc, const Plane &pn, Plane &pf,
                    const double (&hc)[2], double (&hn)[2]) {
        if (z + 1 < z1) {
            load_plane(pf, z + 2);
            load_halo(hn, z + 1);
        }
        const int par = z & 1;
        s_edge[par][0][wv][lane] = make_double2(pc[0][0], pc[0][1]);
        s_edge[par][1][wv][lane] = make_double2(pc[kTY - 1][0], pc[kTY - 1][1]);
        __syncthreads();
        double hm0 = hc[0], hm1 = hc[1], hp0 = hc[0], hp1 = hc[1];
        if (wv > 0) {
            const double2 e = s_edge[par][1][wv - 1][lane];
            hm0 = e.x;
            hm1 = e.y;
        }
        if (wv < kW - 1) {
            const double2 e = s_edge[par][0][wv + 1][lane];
            hp0 = e.x;
            hp1 = e.y;
        }
        if (!wave_on) return;  // wave-uniform; the barrier above was met
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            const int y = y0 + i;
            if (y >= n) break;  // wave-uniform
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
                if constexpr (RESID) {
                    double bcc = 0.0;
                    if (z > 0) bcc = __builtin_fma(oB, pm[i][p], bcc);
                    if (y > 0) bcc = __builtin_fma(oB, ym, bcc);
                    if (x > 0) bcc = __builtin_fma(oB, xm, bcc);
                    bcc = __builtin_fma(dB, pc[i][p], bcc);
                    if (x < n - 1) bcc = __builtin_fma(oB, xp, bcc);
                    if (y < n - 1) bcc = __builtin_fma(oB, yp, bcc);
                    if (z < n - 1) bcc = __builtin_fma(oB, pn[i][p], bcc);
                    out[p] = resid_of(acc, bcc, lam_c);
                }
            }
            const int64_t r = x0 + (int64_t)n * y + nn * z;
            if constexpr (VEC) {
                if (a0) *reinterpret_cast<double2 *>(Yc + r) = make_double2(out[0], out[1]);
            } else {
                if (a0) Yc[r] = out[0];
                if (a1) Yc[r + 1] = out[1];
            }
        }
    };

    Plane P0, P1, P2, P3;
    double H0[2], H1[2];
    load_plane(P0, z0 - 1);
    load_plane(P1, z0);
    load_plane(P2, z0 + 1);
    load_halo(H0, z0);
    for (int z = z0; z < z1; z += 4) {
        step(z, P0, P1, P2, P3, H0, H1);
        if (z + 1 >= z1) break;
        step(z + 1, P1, P2, P3, P0, H1, H0);
        if (z + 2 >= z1) break;
        step(z + 2, P2, P3, P0, P1, H0, H1);
        if (z + 3 >= z1) break;
        step(z + 3, P3, P0, P1, P2, H1, H0);
    }
}

template <int kTY, int kW>
void launch_stencil7(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
                     int ncols, int zc, hipStream_t s, const SellMat *B = nullptr,
                     const double *lam = nullptr)
{
    const int n = A.stencil_n;
    const int ygroups = (n + kW * kTY - 1) / (kW * kTY);
    const int zchunks = (n + zc - 1) / zc;
    const int64_t blocks = (int64_t)ygroups * zchunks * ncols;
    const bool vec = (n % 2 == 0) && (ldx % 2 == 0) && (ldy % 2 == 0) &&
                     (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
    const double dB = B ? B->stencil_d : 0.0, oB = B ? B->stencil_o : 0.0;
#define GCGE_ST7(V, R)                                                                      \
    k_stencil7<kTY, kW, V, R><<<(unsigned)blocks, 64 * kW, 0, s>>>(                         \
        n, ygroups, zchunks, zc, A.stencil_d, A.stencil_o, dB, oB, lam, X, ldx, Y, ldy)
    if (B) {
        if (vec) GCGE_ST7(true, true); else GCGE_ST7(false, true);
    } else {
        if (vec) GCGE_ST7(true, false); else GCGE_ST7(false, false);
    }
#undef GCGE_ST7
}
// ---------------------------------------------------------------------
// The FP32 BlockPCG product w = A p over its 8-column tiles (element
// (row, c) of tile t at ((t n + row) 8 + c % 8), one 32-byte row), with the
// partial p.w sums, for a constant-coefficient 7-point A without shift.
// A lane holds one grid point (all 8 columns of its row); ceil(n / 64)
// wavefronts hold an x-line, the other wavefronts of the kW-wavefront
// workgroup further y-segments of kTY lines; each marches its lines along
// z with planes z - 1, z, z + 1 in registers.  The x +- 1 rows come from
// the adjacent lanes (cross-lane moves; the two edge lanes of a wavefront
// load theirs), the far neighbours (y +- 1, z +- 1) from registers and
// LDS, except the workgroup's two outer halo lines.  Each output is summed
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

// Loads run one plane ahead as in k_stencil7.  Wavefronts xw apart in the
// workgroup hold adjacent y-segments of the same x-range and exchange their
// edge lines through LDS (two slots by plane parity, one barrier per plane
// step).
template <int kTY, int kW>
__global__ __launch_bounds__(64 * kW) void k_stencil_pw(int n, int xw, int ygroups,
                                                         int zchunks, int zc, float d, float o,
                                                         const int *__restrict__ tiles,
                                                         const int *__restrict__ ntact_p,
                                                         const float *__restrict__ P,
                                                         float *__restrict__ Wv,
                                                         float *__restrict__ partial,
                                                         int nparts)
{
    constexpr int kE = kTY == 1 ? 1 : 2;  // edge lines per wavefront
    __shared__ f8s s_edge[2][kE][kW][64];  // [parity][first, last][wavefront][lane]
    __shared__ float red[kW * 8];
    const int k = blockIdx.y;
    // both reads issued together (tiles holds at least gridDim.y entries)
    const int tile = tiles[k];
    if (k >= *ntact_p) return;  // workgroup-uniform
    const int64_t L = row_block<true>();
    const int g = (int)(L % ygroups);
    const int q = (int)(L / ygroups);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int segs = kW / xw;
    const int seg = wv / xw;
    const int y0 = (g * segs + seg) * kTY;
    const int x = (wv % xw) * 64 + lane;
    const bool livex = x < n;
    const bool wave_on = y0 < n;  // wave-uniform
    const int z0 = q * zc;
    const int z1 = z0 + zc < n ? z0 + zc : n;
    const int64_t nn = (int64_t)n * n;
    const int64_t N = nn * n;
    const float *Pt = P + (int64_t)tile * N * 8;
    float *Wt = Wv + (int64_t)tile * N * 8;
    const int yh = seg == 0 ? y0 - 1 : y0 + kTY;
    const bool outer = wave_on && (seg == 0 || seg == segs - 1);

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
    using Plane = f8s[kTY];
    auto load_plane = [&](Plane &p, int z) {
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            if (wave_on) {
                p[i] = load(0, y0 + i, z);
            } else {
#pragma unroll
                for (int t = 0; t < 8; ++t) p[i].v[t] = 0.f;
            }
        }
    };
    auto load_halo = [&](f8s &h, int z) {
        if (outer) {
            h = load(0, yh, z);
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) h.v[t] = 0.f;
        }
    };

    float dot[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) dot[t] = 0.f;
    auto step = [&](int z, const Plane &pm, const Plane &pc, const Plane &pn, Plane &pf,
                    const f8s &hc, f8s &hn) {
        if (z + 1 < z1) {
            load_plane(pf, z + 2);
            load_halo(hn, z + 1);
        }
        const int par = z & 1;
        s_edge[par][0][wv][lane] = pc[0];
        if constexpr (kE == 2) s_edge[par][1][wv][lane] = pc[kTY - 1];
        __syncthreads();
        const f8s hm = seg > 0 ? s_edge[par][kE - 1][wv - xw][lane] : hc;
        const f8s hp = seg < segs - 1 ? s_edge[par][0][wv + xw][lane] : hc;
        if (!wave_on) return;  // wave-uniform; the barrier above was met
#pragma unroll
        for (int i = 0; i < kTY; ++i) {
            const int y = y0 + i;
            if (y >= n) break;  // wave-uniform
            // x - 1 is the lane before's point, x + 1 the lane after's;
            // lanes 0 and 63 take theirs from memory (the neighbouring
            // wavefront's points, or zeros off the grid)
            f8s left, right;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                left.v[t] = __shfl_up(pc[i].v[t], 1, 64);
                right.v[t] = __shfl_down(pc[i].v[t], 1, 64);
            }
            if (lane == 0) left = load(-1, y, z);
            if (lane == 63) right = load(1, y, z);
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
            if (livex) *reinterpret_cast<f8s *>(Wt + (x + (int64_t)n * y + nn * z) * 8) = w;
        }
    };

    Plane P0, P1, P2, P3;
    f8s H0, H1;
    load_plane(P0, z0 - 1);
    load_plane(P1, z0);
    load_plane(P2, z0 + 1);
    load_halo(H0, z0);
    for (int z = z0; z < z1; z += 4) {
        step(z, P0, P1, P2, P3, H0, H1);
        if (z + 1 >= z1) break;
        step(z + 1, P1, P2, P3, P0, H1, H0);
        if (z + 2 >= z1) break;
        step(z + 2, P2, P3, P0, P1, H0, H1);
        if (z + 3 >= z1) break;
        step(z + 3, P3, P0, P1, P2, H1, H0);
    }
    // the workgroup's sum of its threads' partial dots, in a fixed order
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float v = dot[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wv * 8 + t] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < kW; ++w) s += red[w * 8 + threadIdx.x];
        partial[(int64_t)(tile * 8 + threadIdx.x) * nparts + L] = s;
    }
}

template <int kTY, int kW>
void launch_stencil_pw(const SellMat &A, int zc, const int *tiles, const int *ntact_p,
                        int nt_bound, const float *P, float *W, float *partial, hipStream_t s)
{
    const int n = A.stencil_n;
    const int xw = (n + 63) / 64;
    const int ygroups = (n + (kW / xw) * kTY - 1) / ((kW / xw) * kTY);
    const int zchunks = (n + zc - 1) / zc;
    const int parts = ygroups * zchunks;
    k_stencil_pw<kTY, kW><<<dim3((unsigned)parts, (unsigned)nt_bound), 64 * kW, 0, s>>>(
        n, xw, ygroups, zchunks, zc, (float)A.stencil_d, (float)A.stencil_o, tiles, ntact_p, P,
        W, partial, parts);
}
// Shapes measured best on a 128^3 grid: two lines per wavefront, four
// wavefronts per workgroup; the FP64 product marches two planes per
// workgroup (1 to 120 columns), the FP32 one 16 (or as many as keep 512
// workgroups over the launch's tiles).
constexpr int k7TY = 2, k7W = 4, k7Zc = 2;
constexpr int kPwTY = 2, kPwW = 4;

int pw_ygroups(int n)
{
    const int xw = (n + 63) / 64;
    return (n + (kPwW / xw) * kPwTY - 1) / ((kPwW / xw) * kPwTY);
}

int pw_chunk(int n, int nt)
{
    int zc = 16;
    while (zc > 4 && (int64_t)pw_ygroups(n) * ((n + zc - 1) / zc) * nt < 512) zc /= 2;
    return zc;
}

}  // namespace

int stencil_pw_parts(const SellMat &A, int nt_bound)
{
    const int n = A.stencil_n;
    if (n < 4 || n > kStencilMaxN || nt_bound <= 0) return 0;
    const int zc = pw_chunk(n, nt_bound);
    return pw_ygroups(n) * ((n + zc - 1) / zc);
}

void stencil_pw(const SellMat &A, const int *tiles, const int *ntact_p, int nt_bound,
                const float *P, float *W, float *partial, hipStream_t s)
{
    launch_stencil_pw<kPwTY, kPwW>(A, pw_chunk(A.stencil_n, nt_bound), tiles, ntact_p,
                                   nt_bound, P, W, partial, s);
}

bool spmm_stencil(const SellMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy,
                  int ncols, hipStream_t s)
{
    if (A.stencil_n < 4 || A.stencil_n > kStencilMaxN || ncols <= 0) return false;
    launch_stencil7<k7TY, k7W>(A, X, ldx, Y, ldy, ncols, k7Zc, s);
    return true;
}

bool stencil_residual(const SellMat &A, const SellMat &B, const double *X, int64_t ldx,
                      const double *lam, double *R, int64_t ldr, int ncols, hipStream_t s)
{
    if (A.stencil_n < 4 || A.stencil_n > kStencilMaxN || B.stencil_n != A.stencil_n ||
        ncols <= 0)
        return false;
    launch_stencil7<k7TY, k7W>(A, X, ldx, R, ldr, ncols, k7Zc, s, &B, lam);
    return true;
}

}  // namespace gcge

// Grid operators assembled on the device straight into the SELL-64 layout
// (64-row slices, column-major inside a slice, padded with col = row,
// val = 0), the layout HIPCCS_Create produces from a host CCS.
//
// The reference's callers build their matrices on the host and hand over
// a CCSMAT (app/app_ccs.h:20-24, app/app_ccs.c:40-49); for a grid problem
// that is a host pass over every nonzero and a PCIe copy of 12-16 bytes per
// nonzero.  Here only the slice widths (from the grid geometry, on the
// host) and, for variable coefficients, the N node values travel; each
// thread writes one row of one slice, so the column and value stores of a
// slice's 64 lanes are coalesced.
//
// The arithmetic is written to give the same doubles as the host
// generators: products and sums in the generators' order, correctly
// rounded division, no contraction into fused multiply-adds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "app_hip_internal.h"
#include "gcge_hip.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace gcge {
namespace {

constexpr int kBlock = 256;

struct GridArgs {
    int n = 0;
    int64_t N = 0;
    int64_t nslices = 0;
    const double *coef = nullptr;   // node coefficients or the diagonal
    const double *diag0 = nullptr;  // 27 points: diagonal before the shift
    const double *dmax = nullptr;   // 27 points: largest such diagonal
    double delta_rel = 0.0;
    double dval = 0.0, oval = 0.0;  // constant stencils
    double g[4] = {0.0, 0.0, 0.0, 0.0};  // 27 points: 1 / |d|^2
    const int64_t *sp = nullptr;
    int *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
};

enum Form { kConst5, kConst7, kDiff5, kDiff27, kDiag };

// 2 a b / (a + b) with a the coefficient of the lower-numbered node, as the
// generators evaluate it: ((2 a) b) / (a + b).
__device__ inline double harmonic(double a, double b) { return 2.0 * a * b / (a + b); }

template <Form F>
__global__ void __launch_bounds__(kBlock) k_grid_fill(GridArgs a)
{
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // slice row slot
    const int64_t s = r >> 6;
    if (s >= a.nslices) return;
    const int64_t base = a.sp[s] + (r & 63);
    const int w = (int)((a.sp[s + 1] - a.sp[s]) >> 6);
    int j = 0;
    auto put = [&](int64_t c, double v) {
        const int64_t e = base + (int64_t)j * 64;
        a.col[e] = (int)c;
        a.val[e] = v;
        if (a.val32) a.val32[e] = (float)v;
        ++j;
    };
    if (r < a.N) {
        const int64_t n = a.n;
        if constexpr (F == kConst5) {
            const int64_t x = r % n, y = r / n;
            if (y > 0) put(r - n, a.oval);
            if (x > 0) put(r - 1, a.oval);
            put(r, a.dval);
            if (x < n - 1) put(r + 1, a.oval);
            if (y < n - 1) put(r + n, a.oval);
        } else if constexpr (F == kConst7) {
            const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
            if (z > 0) put(r - n * n, a.oval);
            if (y > 0) put(r - n, a.oval);
            if (x > 0) put(r - 1, a.oval);
            put(r, a.dval);
            if (x < n - 1) put(r + 1, a.oval);
            if (y < n - 1) put(r + n, a.oval);
            if (z < n - 1) put(r + n * n, a.oval);
        } else if constexpr (F == kDiff5) {
            const int64_t x = r % n, y = r / n;
            const double kr = a.coef[r];
            const double wxm = x > 0 ? harmonic(a.coef[r - 1], kr) : 0.0;
            const double wxp = x < n - 1 ? harmonic(kr, a.coef[r + 1]) : 0.0;
            const double wym = y > 0 ? harmonic(a.coef[r - n], kr) : 0.0;
            const double wyp = y < n - 1 ? harmonic(kr, a.coef[r + n]) : 0.0;
            // edge weights in the order the edges are visited (+x, -x, +y,
            // -y), then k per missing neighbour
            double d = 0.0;
            if (x < n - 1) d = d + wxp;
            if (x > 0) d = d + wxm;
            if (y < n - 1) d = d + wyp;
            if (y > 0) d = d + wym;
            const int nb = (x == 0) + (x == n - 1) + (y == 0) + (y == n - 1);
            d = d + (double)nb * kr;
            if (y > 0) put(r - n, -wym);
            if (x > 0) put(r - 1, -wxm);
            put(r, d);
            if (x < n - 1) put(r + 1, -wxp);
            if (y < n - 1) put(r + n, -wyp);
        } else if constexpr (F == kDiff27) {
            const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
            const double kr = a.coef[r];
            const double shift = a.delta_rel * a.dmax[0];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dz == 0 && dy == 0 && dx == 0) {
                            put(r, a.diag0[r] + shift);
                            continue;
                        }
                        const bool ok = (dz >= 0 || z > 0) && (dz <= 0 || z < n - 1) &&
                                        (dy >= 0 || y > 0) && (dy <= 0 || y < n - 1) &&
                                        (dx >= 0 || x > 0) && (dx <= 0 || x < n - 1);
                        if (!ok) continue;
                        const int64_t c = r + dz * n * n + dy * n + dx;
                        const double kc = a.coef[c];
                        put(c, -(a.g[dz * dz + dy * dy + dx * dx] * (2.0 * kr * kc / (kr + kc))));
                    }
        } else {
            put(r, a.coef[r]);
        }
    }
    const int64_t pad = r < a.N ? r : 0;
    while (j < w) put(pad, 0.0);
}

// 27 points, first pass: the diagonal before the shift, summed in ascending
// offset order (the weight of a present neighbour, g(d) k_i for a missing
// one), and the largest diagonal of each block.
__global__ void __launch_bounds__(kBlock) k_diag27(GridArgs a, double *diag0, double *bmax)
{
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double d = -INFINITY;
    if (r < a.N) {
        const int64_t n = a.n;
        const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
        const double kr = a.coef[r];
        d = 0.0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dz == 0 && dy == 0 && dx == 0) continue;
                    const double g = a.g[dz * dz + dy * dy + dx * dx];
                    const bool ok = (dz >= 0 || z > 0) && (dz <= 0 || z < n - 1) &&
                                    (dy >= 0 || y > 0) && (dy <= 0 || y < n - 1) &&
                                    (dx >= 0 || x > 0) && (dx <= 0 || x < n - 1);
                    if (ok) {
                        const double kc = a.coef[r + dz * n * n + dy * n + dx];
                        d = d + g * (2.0 * kr * kc / (kr + kc));
                    } else {
                        d = d + g * kr;
                    }
                }
        diag0[r] = d;
    }
    for (int o = 32; o > 0; o >>= 1) d = fmax(d, __shfl_xor(d, o));
    __shared__ double wmax[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int i = 1; i < kBlock / 64; ++i) m = fmax(m, wmax[i]);
        bmax[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(1024) k_max(const double *bmax, int64_t nb, double *out)
{
    double m = -INFINITY;
    for (int64_t i = threadIdx.x; i < nb; i += 1024) m = fmax(m, bmax[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    __shared__ double wmax[16];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 16; ++i) m = fmax(m, wmax[i]);
        out[0] = m;
    }
}

// Row length from the grid position.
int row_len(int form, int64_t r, int64_t n)
{
    const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
    const int ex = (x > 0) + (x < n - 1), ey = (y > 0) + (y < n - 1), ez = (z > 0) + (z < n - 1);
    switch (form) {
    case kConst5:
    case kDiff5: return 1 + ex + ey;
    case kConst7: return 1 + ex + ey + ez;
    case kDiff27: return (1 + ex) * (1 + ey) * (1 + ez);
    default: return 1;
    }
}

template <Form F>
void launch_fill(const GridArgs &a, hipStream_t st)
{
    const int64_t nblk = (a.nslices * 64 + kBlock - 1) / kBlock;
    k_grid_fill<F><<<dim3((unsigned)nblk), dim3(kBlock), 0, st>>>(a);
    HIPCHECK(hipGetLastError());
}

}  // namespace
}  // namespace gcge

using namespace gcge;

extern "C" {

void *HIPCCS_CreateGrid(const gcge_grid_op *op, int flags)
{
    if (op == nullptr || op->n < 1 || op->dim < 1 || op->dim > 3) return nullptr;
    int form;
    switch (op->kind) {
    case GCGE_GRID_LAPLACE:
    case GCGE_GRID_MASS:
        if (op->dim == 2 && op->points == 5)
            form = kConst5;
        else if (op->dim == 3 && op->points == 7)
            form = kConst7;
        else
            return nullptr;
        break;
    case GCGE_GRID_DIFFUSION:
        if (op->dim == 2 && op->points == 5)
            form = kDiff5;
        else if (op->dim == 3 && op->points == 27)
            form = kDiff27;
        else
            return nullptr;
        if (op->coef == nullptr) return nullptr;
        break;
    case GCGE_GRID_DIAGONAL:
        if (op->coef == nullptr) return nullptr;
        form = kDiag;
        break;
    default: return nullptr;
    }
    const int64_t n = op->n;
    int64_t N = 1;
    for (int d = 0; d < op->dim; ++d) N *= n;
    if (N >= ((int64_t)1 << 31)) return nullptr;  // int32 column indices

    Ctx &c = ctx();
    hipStream_t st = c.stream;
    Mat *M = new Mat();
    M->n = N;
    SellMat &S = M->sell;
    S.nrows = N;
    S.nslices = (N + 63) / 64;
    std::vector<int64_t> sp((size_t)S.nslices + 1, 0);
    int maxw = 0;
    int64_t nnz = 0;
    for (int64_t s = 0; s < S.nslices; ++s) {
        int w = 0;
        for (int64_t r = s * 64; r < (s + 1) * 64 && r < N; ++r) {
            const int len = row_len(form, r, n);
            nnz += len;
            if (len > w) w = len;
        }
        if (w > maxw) maxw = w;
        sp[s + 1] = sp[s] + (int64_t)w * 64;
    }
    S.max_width = maxw;
    S.nnz = sp[S.nslices];
    S.nnz_true = nnz;
    S.structure_hash = slice_hash(sp);

    dev_alloc(&S.slice_ptr, sizeof(int64_t) * sp.size());
    dev_alloc(&S.col, sizeof(int) * (size_t)S.nnz);
    dev_alloc(&S.val, sizeof(double) * (size_t)S.nnz);
    M->bytes = (int64_t)(sizeof(int64_t) * sp.size() + (sizeof(int) + sizeof(double)) * S.nnz);
    if (flags & GCGE_HIP_MAT_FP32) {
        dev_alloc(&S.val32, sizeof(float) * (size_t)S.nnz);
        M->bytes += (int64_t)(sizeof(float) * S.nnz);
    }
    HIPCHECK(hipMemcpyAsync(S.slice_ptr, sp.data(), sizeof(int64_t) * sp.size(),
                            hipMemcpyHostToDevice, st));

    GridArgs a;
    a.n = op->n;
    a.N = N;
    a.nslices = S.nslices;
    a.sp = S.slice_ptr;
    a.col = S.col;
    a.val = S.val;
    a.val32 = S.val32;
    const double twod = 2.0 * op->dim;
    if (op->kind == GCGE_GRID_LAPLACE) {
        a.dval = twod;
        a.oval = -1.0;
    } else if (op->kind == GCGE_GRID_MASS) {
        a.dval = 2.0 * twod / (3.0 * twod);
        a.oval = 1.0 / (3.0 * twod);
    }
    a.delta_rel = op->delta_rel;
    a.g[1] = 1.0;
    a.g[2] = 1.0 / 2.0;
    a.g[3] = 1.0 / 3.0;
    // scratch: the coefficients, and for 27 points the unshifted diagonal,
    // the block maxima and the maximum
    const int64_t nblk = (N + kBlock - 1) / kBlock;
    double *tmp = nullptr;
    if (op->coef != nullptr && (form == kDiff5 || form == kDiff27 || form == kDiag)) {
        const int64_t extra = form == kDiff27 ? N + nblk + 1 : 0;
        dev_alloc(&tmp, sizeof(double) * (size_t)(N + extra));
        HIPCHECK(hipMemcpyAsync(tmp, op->coef, sizeof(double) * (size_t)N, hipMemcpyHostToDevice,
                                st));
        a.coef = tmp;
    }
    switch (form) {
    case kConst5: launch_fill<kConst5>(a, st); break;
    case kConst7: launch_fill<kConst7>(a, st); break;
    case kDiff5: launch_fill<kDiff5>(a, st); break;
    case kDiag: launch_fill<kDiag>(a, st); break;
    case kDiff27: {
        double *diag0 = tmp + N, *bmax = tmp + 2 * N, *dmax = tmp + 2 * N + nblk;
        k_diag27<<<dim3((unsigned)nblk), dim3(kBlock), 0, st>>>(a, diag0, bmax);
        HIPCHECK(hipGetLastError());
        k_max<<<1, 1024, 0, st>>>(bmax, nblk, dmax);
        HIPCHECK(hipGetLastError());
        a.diag0 = diag0;
        a.dmax = dmax;
        launch_fill<kDiff27>(a, st);
        break;
    }
    }
    HIPCHECK(hipStreamSynchronize(st));  // the host coefficients and sp may go now
    if (tmp) dev_free(tmp);
    detect_plane(S, st);
    if (form == kConst7 && !(flags & GCGE_HIP_MAT_GENERAL) && n >= 4 && n <= kStencilMaxN) {
        S.stencil_n = (int)n;
        S.stencil_d = a.dval;
        S.stencil_o = a.oval;
    }
    return M;
}

int HIPCCS_Equal(void *pa, void *pb)
{
    if (pa == nullptr || pb == nullptr) return 0;
    Mat *A = mat_of(pa, "HIPCCS_Equal");
    Mat *B = mat_of(pb, "HIPCCS_Equal");
    const SellMat &x = A->sell, &y = B->sell;
    if (A->n != B->n || x.nslices != y.nslices || x.nnz != y.nnz || x.nnz_true != y.nnz_true ||
        x.max_width != y.max_width || x.structure_hash != y.structure_hash)
        return 0;
    ctx().sync();
    auto same = [](const void *p, const void *q, size_t bytes) {
        std::vector<unsigned char> hp(bytes), hq(bytes);
        if (bytes == 0) return true;
        HIPCHECK(hipMemcpy(hp.data(), p, bytes, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hq.data(), q, bytes, hipMemcpyDeviceToHost));
        return hp == hq;
    };
    if (!same(x.slice_ptr, y.slice_ptr, sizeof(int64_t) * (size_t)(x.nslices + 1))) return 0;
    if (!same(x.col, y.col, sizeof(int) * (size_t)x.nnz)) return 0;
    if (!same(x.val, y.val, sizeof(double) * (size_t)x.nnz)) return 0;
    if (x.val32 && y.val32 && !same(x.val32, y.val32, sizeof(float) * (size_t)x.nnz)) return 0;
    return 1;
}

}  // extern "C"

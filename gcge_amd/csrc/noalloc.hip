// Inner products of multivectors without any device allocation.
//
// HIP_MultiVecInnerProdNoAlloc is the table's MultiVecInnerProd slot
// (app/app_lapack.c:299-321 semantics: 'N' full block, 'S' symmetric,
// 'D' diagonal, results in a host array with leading dimension ldIP) for
// callers that cannot call hipMalloc.  The slot keeps its Gram and dot
// partial sums and its device result in buffers the library allocates and
// grows on demand; this variant keeps both in `scratch`, a file-scope
// device array of kScratchDoubles doubles (32 MiB) that every GPU the
// library runs on holds until the process ends, and copies the result
// straight into the caller's array.  It runs the slot's kernels with the
// same launch plan whenever that plan's partial sums fit, so the results
// are the slot's bit for bit; otherwise it groups the rows in as many
// chunks as fit (a different summation grouping), and it never runs the
// launch-shape tuner, which allocates.
//
// Every call on one GPU uses that GPU's one array and leaves its partial
// sums there; each call overwrites what it reads before reading it and
// waits for its stream before returning, so calls made one after another
// are independent, but calls that overlap in time on one GPU (from
// different host threads) race on the array.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "app_hip_internal.h"
#include "gcge_hip.h"
#include "kernels.h"

namespace {

constexpr int64_t kScratchDoubles = (int64_t)4 << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

// scratch's device address on each GPU, looked up once per device with
// that device current (the array's host-side name is not a device address)
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {nullptr};

double *scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            char buf[200];
            snprintf(buf, sizeof(buf), "HIP_MultiVecInnerProdNoAlloc: hipGetSymbolAddress: %s",
                     hipGetErrorString(e));
            gcge::record_error(buf);
            return nullptr;
        }
        g_ptr[dev] = static_cast<double *>(p);
    }
    return g_ptr[dev];
}

int reject(const char *what)
{
    char buf[256];
    snprintf(buf, sizeof(buf), "HIP_MultiVecInnerProdNoAlloc: %s", what);
    gcge::record_error(buf);
    return -1;
}

}  // namespace

extern "C" int HIP_MultiVecInnerProdNoAlloc(char nsdIP, void **x, void **y, int is_vec,
                                            int *start, int *end, double *inner_prod,
                                            int ldIP)
{
    using namespace gcge;
    (void)is_vec;
    MV *X = mv_of(x, "MultiVecInnerProdNoAlloc x");
    MV *Y = mv_of(y, "MultiVecInnerProdNoAlloc y");
    const int nr = end[0] - start[0], nc = end[1] - start[1];
    if (nr <= 0 || nc <= 0) return 0;
    if (start[0] < 0 || end[0] > X->ncols || start[1] < 0 || end[1] > Y->ncols)
        return reject("column range outside the multivector");
    if (X->nrows != Y->nrows) return reject("row counts differ");
    if (nsdIP != 'N' && nsdIP != 'S' && nsdIP != 'D') return reject("nsdIP is not N, S or D");
    if ((nsdIP == 'S' || nsdIP == 'D') && nr != nc) return reject("'S'/'D' need equal ranges");
    if (nsdIP != 'D' && ldIP < nr) return reject("ldIP below the row range");
    Ctx &c = ctx();
    double *base = scratch_on(c.device);
    if (base == nullptr) return -1;
    const double *xs = X->d + (int64_t)start[0] * X->ldd;
    const double *ys = Y->d + (int64_t)start[1] * Y->ldd;
    hipStream_t st = c.stream;
    if (nsdIP == 'D') {
        const int64_t part = dots_scratch(nc);
        const int64_t res_off = (part + 31) / 32 * 32;
        if (res_off + nc > kScratchDoubles)
            return reject("more columns than the 32 MiB scratch array holds");
        col_dots(X->nrows, nc, xs, X->ldd, ys, Y->ldd, base + res_off, base, st);
        HIPCHECK(hipMemcpy2DAsync(inner_prod, sizeof(double) * (size_t)ldIP, base + res_off,
                                  sizeof(double), sizeof(double), (size_t)nc,
                                  hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return 0;
    }
    const bool sym = nsdIP == 'S';
    const int64_t res = (int64_t)nr * nc;
    if (res > kScratchDoubles) return reject("result block larger than the 32 MiB scratch array");
    const GramPlan plan =
        gram_plan_bounded(X->nrows, nr, nc, sym, kScratchDoubles - ((res + 31) / 32 * 32));
    if (plan.scratch <= 0) return reject("partial sums do not fit the 32 MiB scratch array");
    const int64_t res_off = (plan.scratch + 31) / 32 * 32;
    double *G = base + res_off;
    gram(plan, X->nrows, nr, nc, xs, X->ldd, ys, Y->ldd, G, nr, sym, base, st);
    HIPCHECK(hipMemcpy2DAsync(inner_prod, sizeof(double) * (size_t)ldIP, G,
                              sizeof(double) * (size_t)nr, sizeof(double) * (size_t)nr,
                              (size_t)nc, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (sym) {
        // the kernel computed the lower triangle; mirror it
        for (int j = 0; j < nc; ++j)
            for (int i = j + 1; i < nr; ++i)
                inner_prod[j + (int64_t)ldIP * i] = inner_prod[i + (int64_t)ldIP * j];
    }
    return 0;
}

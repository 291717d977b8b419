// The HIP "app" of the GCG eigensolver: fills the operation table with
// device implementations (the role app/app_ccs.c:213-249 plays for the
// CPU).  Multivectors and matrices live in device memory; only the small
// coefficient and result matrices cross to the host, as the table's
// signatures require (src/ops.h:78-103).
//
// Slot semantics follow app/app_ccs.c:50-150 (sparse products),
// app/app_lapack.c:230-534 (multivector algebra) and
// src/ops_multi_vec.c:351-411 (MultiVecQtAP).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "app_hip_internal.h"
#include "gcge_hip.h"
#include "dense_eig.h"
#include "kernels.h"
#include "tuner.h"

namespace gcge {

// ------------------------------------------------------------------ errors
static std::mutex g_err_mu;
static std::string g_err;
static int g_err_code = 0;

void fail(const char *what, const char *file, int line)
{
    {
        std::lock_guard<std::mutex> lk(g_err_mu);
        char buf[512];
        snprintf(buf, sizeof(buf), "gcge_hip: %s (%s:%d)", what, file, line);
        g_err = buf;
        g_err_code = 1;
    }
    fprintf(stderr, "%s\n", g_err.c_str());
    fflush(stderr);
    abort();
}

void record_error(const char *what)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = std::string("gcge_hip: ") + what;
    g_err_code = 1;
}

void hip_check(hipError_t e, const char *expr, const char *file, int line)
{
    if (e == hipSuccess) return;
    char buf[400];
    snprintf(buf, sizeof(buf), "%s failed: %s", expr, hipGetErrorString(e));
    fail(buf, file, line);
}

// ----------------------------------------------------------------- context
namespace {
thread_local Ctx *t_ctx = nullptr;
std::mutex g_ctx_mu;  // the registry of live contexts
int g_live_ctx = 0;
// Kernel timing of every thread: contexts created while it is on start
// with their profiler on, and released contexts add their totals here.
std::mutex g_prof_mu;
bool g_prof_all = false;
int64_t g_prof_launches[PROF_NKIND] = {0};
double g_prof_ms[PROF_NKIND] = {0}, g_prof_bytes[PROF_NKIND] = {0}, g_prof_flops[PROF_NKIND] = {0};
}

Ctx &ctx()
{
    if (t_ctx == nullptr) {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        tune_load_default();
        Ctx *x = new Ctx();
        {
            std::lock_guard<std::mutex> lp(g_prof_mu);
            x->prof.on = g_prof_all;
        }
        HIPCHECK(hipGetDevice(&x->device));
        HIPCHECK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
        t_ctx = x;
        ++g_live_ctx;
    }
    return *t_ctx;
}

// ------------------------------------------------------------- memory pool
struct MemPool {
    bool on = false;
    std::multimap<size_t, void *> cached_dev, cached_host;
    size_t cached_bytes = 0;
};

namespace {
// Cached bytes a pool keeps at most; past it the largest cached blocks go
// back to HIP.
constexpr size_t kPoolCap = (size_t)16 << 30;
struct Block {
    MemPool *pool;
    size_t size;
    bool host;
};
std::mutex g_blk_mu;                         // g_blocks and every pool's lists
std::unordered_map<void *, Block> g_blocks;  // blocks allocated under a pool

void *raw_alloc(size_t bytes, bool host)
{
    void *p = nullptr;
    if (host)
        HIPCHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    else
        HIPCHECK(hipMalloc(&p, bytes));
    return p;
}

void raw_free(void *p, bool host)
{
    if (host)
        HIPCHECK(hipHostFree(p));
    else
        HIPCHECK(hipFree(p));
}

void *pool_alloc(size_t bytes, bool host)
{
    if (bytes == 0) bytes = 1;
    Ctx &c = ctx();
    MemPool *pool = c.pool;
    if (pool == nullptr || !pool->on) return raw_alloc(bytes, host);
    {
        std::lock_guard<std::mutex> lk(g_blk_mu);
        auto &lst = host ? pool->cached_host : pool->cached_dev;
        auto it = lst.lower_bound(bytes);
        if (it != lst.end() && it->first <= 2 * bytes + 65536) {
            void *p = it->second;
            pool->cached_bytes -= it->first;
            lst.erase(it);
            return p;
        }
    }
    const size_t size = (bytes + 255) / 256 * 256;
    void *p = raw_alloc(size, host);
    std::lock_guard<std::mutex> lk(g_blk_mu);
    g_blocks[p] = Block{pool, size, host};
    return p;
}

void pool_free(void *p, bool host)
{
    if (p == nullptr) return;
    std::vector<std::pair<void *, bool>> release;
    {
        std::lock_guard<std::mutex> lk(g_blk_mu);
        auto it = g_blocks.find(p);
        if (it != g_blocks.end()) {
            const Block b = it->second;
            if (b.pool != nullptr && b.pool->on) {
                (b.host ? b.pool->cached_host : b.pool->cached_dev).insert({b.size, p});
                b.pool->cached_bytes += b.size;
                // over the cap: hand the largest cached blocks back
                while (b.pool->cached_bytes > kPoolCap) {
                    auto &ld = b.pool->cached_dev, &lh = b.pool->cached_host;
                    const bool from_dev =
                        !ld.empty() && (lh.empty() || ld.rbegin()->first >= lh.rbegin()->first);
                    auto &lst = from_dev ? ld : lh;
                    auto last = std::prev(lst.end());
                    b.pool->cached_bytes -= last->first;
                    release.push_back({last->second, !from_dev});
                    g_blocks.erase(last->second);
                    lst.erase(last);
                }
                if (release.empty()) return;
                p = nullptr;
            } else {
                g_blocks.erase(it);
            }
        }
    }
    for (auto &r : release) raw_free(r.first, r.second);
    if (p != nullptr) raw_free(p, host);
}

// Hands a pool's cached blocks back to HIP and turns it off; blocks still in
// use are freed to HIP when they are freed.
void pool_flush(MemPool *pool, bool forget)
{
    std::vector<std::pair<void *, bool>> release;
    {
        std::lock_guard<std::mutex> lk(g_blk_mu);
        pool->on = false;
        for (auto *lst : {&pool->cached_dev, &pool->cached_host}) {
            for (auto &e : *lst) {
                release.push_back({e.second, lst == &pool->cached_host});
                g_blocks.erase(e.second);
            }
            lst->clear();
        }
        pool->cached_bytes = 0;
        if (forget)
            for (auto &e : g_blocks)
                if (e.second.pool == pool) e.second.pool = nullptr;
    }
    for (auto &r : release) raw_free(r.first, r.second);
}
}  // namespace

void *dev_alloc_bytes(size_t bytes) { return pool_alloc(bytes, false); }
void *pinned_alloc_bytes(size_t bytes) { return pool_alloc(bytes, true); }
void dev_free(void *p) { pool_free(p, false); }
void pinned_free(void *p) { pool_free(p, true); }

uint64_t slice_hash(const std::vector<int64_t> &sp)
{
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the slice pointers
    for (int64_t v : sp) {
        h ^= (uint64_t)v;
        h *= 1099511628211ull;
    }
    return h;
}

std::unique_lock<std::mutex> sole_context_lock()
{
    std::unique_lock<std::mutex> lk(g_ctx_mu);
    if (t_ctx == nullptr || g_live_ctx != 1) lk.unlock();
    return lk;
}

void release_ctx()
{
    Ctx *c = t_ctx;
    if (c == nullptr) return;
    c->sync();
    c->prof.collect();
    {
        std::lock_guard<std::mutex> lp(g_prof_mu);
        if (g_prof_all)
            for (int k = 0; k < PROF_NKIND; ++k) {
                g_prof_launches[k] += c->prof.launches[k];
                g_prof_ms[k] += c->prof.ms[k];
                g_prof_bytes[k] += c->prof.bytes[k];
                g_prof_flops[k] += c->prof.flops[k];
            }
    }
    if (c->dev_eig) release_dev_eig(c->dev_eig);
    if (c->d_scratch) dev_free(c->d_scratch);
    if (c->d_small) dev_free(c->d_small);
    if (c->h_pinned) pinned_free(c->h_pinned);
    for (Ctx::Staging &g : c->stage) {
        if (g.h) pinned_free(g.h);
        if (g.d) dev_free(g.d);
        if (g.ev) HIPCHECK(hipEventDestroy(g.ev));
    }
    for (hipEvent_t e : c->prof.pool) HIPCHECK(hipEventDestroy(e));
    if (c->pool) {
        pool_flush(c->pool, true);
        delete c->pool;
    }
    HIPCHECK(hipStreamDestroy(c->stream));
    delete c;
    t_ctx = nullptr;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    --g_live_ctx;
}

double *Ctx::scratch(int64_t ndouble)
{
    if (ndouble > scratch_cap) {
        sync();
        if (d_scratch) dev_free(d_scratch);
        int64_t cap = ndouble + ndouble / 4 + 1024;
        dev_alloc(&d_scratch, cap * sizeof(double));
        scratch_cap = cap;
    }
    return d_scratch;
}

double *Ctx::small(int64_t ndouble)
{
    if (ndouble > small_cap) {
        sync();
        if (d_small) dev_free(d_small);
        int64_t cap = ndouble + 4096;
        dev_alloc(&d_small, cap * sizeof(double));
        small_cap = cap;
    }
    return d_small;
}

double *Ctx::pinned(int64_t ndouble)
{
    if (ndouble > pinned_cap) {
        // the stream may still read the old staging buffer
        HIPCHECK(hipStreamSynchronize(stream));
        if (h_pinned) pinned_free(h_pinned);
        int64_t cap = ndouble + 4096;
        pinned_alloc(&h_pinned, cap * sizeof(double));
        pinned_cap = cap;
    }
    return h_pinned;
}

void Ctx::sync() { HIPCHECK(hipStreamSynchronize(stream)); }

Ctx::Staging &Ctx::staging(int64_t ndouble)
{
    Staging &g = stage[stage_next];
    stage_next = (stage_next + 1) % kStage;
    if (g.pending) {
        HIPCHECK(hipEventSynchronize(g.ev));
        g.pending = false;
    }
    if (ndouble > g.cap) {
        if (g.h) pinned_free(g.h);
        if (g.d) dev_free(g.d);
        g.cap = ndouble + 4096;
        pinned_alloc(&g.h, g.cap * sizeof(double));
        dev_alloc(&g.d, g.cap * sizeof(double));
    }
    if (g.ev == nullptr) HIPCHECK(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
    return g;
}

void Ctx::staged(Staging &g)
{
    HIPCHECK(hipEventRecord(g.ev, stream));
    g.pending = true;
}

// ---------------------------------------------------------------- profiler
hipEvent_t Profiler::begin(int kind, double b, double f, hipStream_t s)
{
    if (!on) return nullptr;
    auto take = [&]() {
        hipEvent_t e;
        if (!pool.empty()) {
            e = pool.back();
            pool.pop_back();
        } else {
            HIPCHECK(hipEventCreate(&e));
        }
        return e;
    };
    Rec r;
    r.a = take();
    r.b = take();
    r.kind = kind;
    r.bytes = b;
    r.flops = f;
    HIPCHECK(hipEventRecord(r.a, s));
    pending.push_back(r);
    return r.b;
}

void Profiler::end(hipEvent_t e, hipStream_t s)
{
    if (e) HIPCHECK(hipEventRecord(e, s));
    if (pending.size() > 4096) collect();
}

void Profiler::collect()
{
    for (Rec &r : pending) {
        HIPCHECK(hipEventSynchronize(r.b));
        float t = 0.f;
        HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
        launches[r.kind] += 1;
        ms[r.kind] += t;
        bytes[r.kind] += r.bytes;
        flops[r.kind] += r.flops;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}

void Profiler::reset()
{
    collect();
    for (int k = 0; k < PROF_NKIND; ++k) {
        launches[k] = 0;
        ms[k] = bytes[k] = flops[k] = 0.0;
    }
}

// ------------------------------------------------------------- multivector
MV *mv_of(void **p, const char *where)
{
    MV *m = reinterpret_cast<MV *>(p);
    if (m == nullptr || m->magic != MV::kMagic) {
        char buf[128];
        snprintf(buf, sizeof(buf), "%s: not a HIP multivector handle", where);
        fail(buf, __FILE__, __LINE__);
    }
    return m;
}

Mat *mat_of(void *p, const char *where)
{
    Mat *m = static_cast<Mat *>(p);
    if (m == nullptr || m->magic != Mat::kMagic) {
        char buf[128];
        snprintf(buf, sizeof(buf), "%s: not a HIP matrix handle", where);
        fail(buf, __FILE__, __LINE__);
    }
    return m;
}

MV *mv_create(int64_t nrows, int64_t ncols)
{
    Ctx &c = ctx();
    MV *m = new MV();
    m->nrows = nrows;
    m->ncols = ncols;
    m->ldd = nrows;
    m->owner = true;
    const size_t bytes = (size_t)(nrows * (ncols > 0 ? ncols : 1)) * sizeof(double);
    dev_alloc(&m->d, bytes);
    HIPCHECK(hipMemsetAsync(m->d, 0, bytes, c.stream));
    // another thread's stream may use the multivector next
    c.sync();
    return m;
}

void mv_destroy(MV *m)
{
    if (!m) return;
    if (m->owner && m->d) {
        ctx().sync();
        dev_free(m->d);
    }
    m->magic = 0;
    delete m;
}

static void check_range(const MV *m, int s, int e, const char *where)
{
    if (s < 0 || e < s || e > m->ncols) {
        char buf[160];
        snprintf(buf, sizeof(buf), "%s: column range [%d,%d) outside [0,%lld)", where, s, e,
                 (long long)m->ncols);
        fail(buf, __FILE__, __LINE__);
    }
}

}  // namespace gcge

using namespace gcge;

namespace {

// ------------------------------------------------------------------- slots
void h_MultiVecCreateByMat(void ***des, int num_vec, void *src_mat, OPS *)
{
    Mat *A = mat_of(src_mat, "MultiVecCreateByMat");
    *des = reinterpret_cast<void **>(mv_create(A->n, num_vec));
}

void h_MultiVecCreateByVec(void ***des, int num_vec, void *src_vec, OPS *)
{
    MV *s = mv_of(reinterpret_cast<void **>(src_vec), "MultiVecCreateByVec");
    *des = reinterpret_cast<void **>(mv_create(s->nrows, num_vec));
}

void h_MultiVecCreateByMultiVec(void ***des, int num_vec, void **src, OPS *)
{
    MV *s = mv_of(src, "MultiVecCreateByMultiVec");
    *des = reinterpret_cast<void **>(mv_create(s->nrows, num_vec));
}

void h_MultiVecDestroy(void ***des, int, OPS *)
{
    if (*des) mv_destroy(mv_of(*des, "MultiVecDestroy"));
    *des = nullptr;
}

void h_GetVecFromMultiVec(void **multi_vec, int c, void **vec, OPS *)
{
    MV *m = mv_of(multi_vec, "GetVecFromMultiVec");
    check_range(m, c, c + 1, "GetVecFromMultiVec");
    MV *v = new MV();
    v->nrows = m->nrows;
    v->ncols = 1;
    v->ldd = m->ldd;
    v->d = m->d + (int64_t)c * m->ldd;
    v->owner = false;
    *vec = v;
}

void h_RestoreVecForMultiVec(void **, int, void **vec, OPS *)
{
    if (*vec) mv_destroy(mv_of(reinterpret_cast<void **>(*vec), "RestoreVecForMultiVec"));
    *vec = nullptr;
}

void h_MultiVecView(void **x, int start, int end, OPS *ops)
{
    MV *m = mv_of(x, "MultiVecView");
    check_range(m, start, end, "MultiVecView");
    const int nc = end - start;
    std::vector<double> h((size_t)m->nrows * (nc > 0 ? nc : 1));
    Ctx &c = ctx();
    for (int k = 0; k < nc; ++k)
        HIPCHECK(hipMemcpyAsync(h.data() + (size_t)m->nrows * k,
                                m->d + (int64_t)(start + k) * m->ldd,
                                sizeof(double) * m->nrows, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int64_t r = 0; r < m->nrows; ++r) {
        for (int k = 0; k < nc; ++k) ops->Printf("%6.4e\t", h[(size_t)m->nrows * k + r]);
        ops->Printf("\n");
    }
}

// y[:, s1:e1] = mat * x[:, s0:e0]; mat NULL copies (app/app_ccs.c:50-139)
void h_MatDotMultiVec(void *mat, void **x, void **y, int *start, int *end, OPS *)
{
    MV *X = mv_of(x, "MatDotMultiVec x"), *Y = mv_of(y, "MatDotMultiVec y");
    const int nc = end[0] - start[0];
    if (nc != end[1] - start[1]) fail("MatDotMultiVec: ranges differ in length", __FILE__, __LINE__);
    check_range(X, start[0], end[0], "MatDotMultiVec x");
    check_range(Y, start[1], end[1], "MatDotMultiVec y");
    if (nc == 0) return;
    Ctx &c = ctx();
    const double *xs = X->d + (int64_t)start[0] * X->ldd;
    double *ys = Y->d + (int64_t)start[1] * Y->ldd;
    if (mat == nullptr) {
        copy_cols(Y->nrows, nc, xs, X->ldd, ys, Y->ldd, c.stream);
        return;
    }
    Mat *A = mat_of(mat, "MatDotMultiVec");
    if (A->n != X->nrows || A->n != Y->nrows)
        fail("MatDotMultiVec: matrix and vector sizes differ", __FILE__, __LINE__);
    // algorithmic bytes: CCS arrays once (none for a matrix-free stencil
    // operator), X read once, Y written once
    const double mat_bytes = A->sell.stencil_n > 0 ? 0.0
                             : (double)A->sell.nnz_true * 12.0 + 4.0 * (double)(A->n + 1);
    const double bytes = mat_bytes + 2.0 * 8.0 * (double)nc * (double)A->n;
    hipEvent_t ev = c.prof.begin(PROF_SPMM64, bytes, 2.0 * (double)A->sell.nnz_true * nc,
                                 c.stream);
    spmm(A->sell, xs, X->ldd, ys, Y->ldd, nc, c.stream);
    c.prof.end(ev, c.stream);
    c.stats.spmm_calls++;
    c.stats.spmm_cols += nc;
}

// r[:, sr:sr+nc] = A x[:, sx:sx+nc] - B x[:, sx:sx+nc] diag(lam) in one pass
// when ops is this table and A, B are stencil operators of one grid; the
// same bits as MatDotMultiVec by A and by B, MultiVecLinearComb scaling by
// lam and MultiVecAxpby(-1, ., 1, .), the residual sequence of GCG's
// convergence check (src/ops_eig_sol_gcg.c:210-222).  False (nothing done)
// otherwise.
bool stencil_residual_slot(OPS *ops, void *matA, void *matB, void **x, int sx, int nc,
                           const double *lam, void **r, int sr)
{
    if (ops == nullptr || ops->MatDotMultiVec != h_MatDotMultiVec || matA == nullptr ||
        matB == nullptr || nc <= 0)
        return false;
    Mat *A = mat_of(matA, "stencil residual A"), *B = mat_of(matB, "stencil residual B");
    if (A->sell.stencil_n == 0 || B->sell.stencil_n != A->sell.stencil_n) return false;
    MV *X = mv_of(x, "stencil residual x"), *R = mv_of(r, "stencil residual r");
    check_range(X, sx, sx + nc, "stencil residual x");
    check_range(R, sr, sr + nc, "stencil residual r");
    if (A->n != X->nrows || A->n != R->nrows || B->n != A->n)
        fail("stencil residual: matrix and vector sizes differ", __FILE__, __LINE__);
    Ctx &c = ctx();
    Ctx::Staging &g = c.staging(nc);
    for (int k = 0; k < nc; ++k) g.h[k] = lam[k];
    HIPCHECK(hipMemcpyAsync(g.d, g.h, sizeof(double) * nc, hipMemcpyHostToDevice, c.stream));
    // X read once, R written once; two products' flops
    hipEvent_t ev = c.prof.begin(PROF_SPMM64, 2.0 * 8.0 * (double)nc * (double)A->n,
                                 2.0 * (double)(A->sell.nnz_true + B->sell.nnz_true) * nc,
                                 c.stream);
    const bool ok = stencil_residual(A->sell, B->sell, X->d + (int64_t)sx * X->ldd, X->ldd, g.d,
                                     R->d + (int64_t)sr * R->ldd, R->ldd, nc, c.stream);
    c.prof.end(ev, c.stream);
    c.staged(g);
    if (!ok) fail("stencil residual: operators not handled", __FILE__, __LINE__);
    c.stats.spmm_calls++;
    c.stats.spmm_cols += nc;
    return true;
}

// y = alpha x + beta y over column ranges (app/app_lapack.c:334-395)
void h_MultiVecAxpby(double alpha, void **x, double beta, void **y, int *start, int *end,
                     OPS *)
{
    MV *Y = mv_of(y, "MultiVecAxpby y");
    const int nc = end[1] - start[1];
    if (nc != end[0] - start[0]) fail("MultiVecAxpby: ranges differ in length", __FILE__, __LINE__);
    check_range(Y, start[1], end[1], "MultiVecAxpby y");
    if (nc == 0 || Y->nrows == 0) return;
    const double *xs = nullptr;
    int64_t ldx = 0;
    if (x != nullptr) {
        MV *X = mv_of(x, "MultiVecAxpby x");
        check_range(X, start[0], end[0], "MultiVecAxpby x");
        if (X->nrows != Y->nrows) fail("MultiVecAxpby: row counts differ", __FILE__, __LINE__);
        xs = X->d + (int64_t)start[0] * X->ldd;
        ldx = X->ldd;
    }
    axpby(Y->nrows, nc, alpha, xs, ldx, beta, Y->d + (int64_t)start[1] * Y->ldd, Y->ldd,
          ctx().stream);
}

// Whether X[:, nr] C + Y diag(beta) (nc destination columns) runs as the
// library GEMM: from 448 Ki rows with at least 100 source and 100
// destination columns and no per-column beta (X C + beta Y is a plain
// GEMM; 47-66 TFLOP/s against 37-53 for the matrix-core kernel at 0.5-4 M
// rows, 150-460 source and 100-220 destination columns); otherwise the
// matrix-core kernel (equal at 262,144-373,248 rows or 320 destination
// columns, faster below; 35 against 23 TFLOP/s for 40 destination columns
// at 4 M rows, whose 128-wide tiles the library fills to a third).
bool lincomb_library(const MV *X, const MV *Y, int nr, int nc, bool per_column_beta)
{
    return Y->nrows >= (int64_t(448) << 10) && nr >= 100 && nc >= 100 && !per_column_beta &&
           Y->nrows <= INT32_MAX && X->ldd <= INT32_MAX && Y->ldd <= INT32_MAX;
}

// Y (nrows x nc) = X (nrows x nr) C + Y diag(beta) with C on the device:
// column-major (nr rows) for the library GEMM, transposed and k-major
// (C[i * nc + k]) for the matrix-core kernel.
void lincomb_launch(Ctx &c, const double *X, int64_t ldx, int64_t nrows, int nr, int nc,
                    const double *dC, bool library, int beta_mode, double beta_scalar,
                    const double *dbeta, double *Y, int64_t ldy)
{
    hipEvent_t ev = c.prof.begin(library ? PROF_GEMM : PROF_LINCOMB,
                                 8.0 * (double)nrows * (nr + nc * (beta_mode ? 2.0 : 1.0)),
                                 2.0 * (double)nrows * nr * nc, c.stream);
    if (library)
        gemm_nn(nrows, nc, nr, X, ldx, dC, nr, beta_mode == 1 ? beta_scalar : 0.0, Y, ldy);
    else
        lincomb(nrows, nr, nc, X, ldx, dC, nc, beta_mode, beta_scalar, dbeta, Y, ldy, c.stream);
    c.prof.end(ev, c.stream);
    c.stats.lincomb_calls++;
    c.stats.lincomb_flops += 2.0 * (double)nrows * nr * nc;
}

// y = x coef + y diag(beta) (app/app_lapack.c:463-534)
void h_MultiVecLinearComb(void **x, void **y, int, int *start, int *end, double *coef, int ldc,
                          double *beta, int incb, OPS *)
{
    MV *Y = mv_of(y, "MultiVecLinearComb y");
    const int nr = end[0] - start[0], nc = end[1] - start[1];
    if (nr == 0 || nc == 0 || Y->nrows == 0) return;
    check_range(Y, start[1], end[1], "MultiVecLinearComb y");
    Ctx &c = ctx();
    double *ys = Y->d + (int64_t)start[1] * Y->ldd;
    const bool have_x = (x != nullptr && coef != nullptr);
    if (!have_x) {
        if (beta == nullptr) return;
        if (incb == 0) {
            if (*beta != 1.0) scale_cols(Y->nrows, nc, nullptr, *beta, ys, Y->ldd, c.stream);
        } else {
            Ctx::Staging &g = c.staging(nc);
            for (int k = 0; k < nc; ++k) g.h[k] = beta[(int64_t)k * incb];
            HIPCHECK(hipMemcpyAsync(g.d, g.h, sizeof(double) * nc, hipMemcpyHostToDevice,
                                    c.stream));
            scale_cols(Y->nrows, nc, g.d, 0.0, ys, Y->ldd, c.stream);
            c.staged(g);
        }
        return;
    }
    MV *X = mv_of(x, "MultiVecLinearComb x");
    check_range(X, start[0], end[0], "MultiVecLinearComb x");
    if (X->nrows != Y->nrows) fail("MultiVecLinearComb: row counts differ", __FILE__, __LINE__);
    // The library takes coef as it is (column-major, nr rows); the kernel
    // transposed and k-major (Ct[i * nc + k] = coef[i + k * ldc]).
    // Coefficients and beta go in one upload.
    const bool library = lincomb_library(X, Y, nr, nc, beta != nullptr && incb != 0);
    const int64_t ncoef = (int64_t)nr * nc;
    Ctx::Staging &g = c.staging(ncoef + nc);
    double *h = g.h;
    for (int k = 0; k < nc; ++k) {
        const double *src = coef + (int64_t)ldc * k;
        if (library)
            memcpy(h + (int64_t)nr * k, src, sizeof(double) * nr);
        else
            for (int i = 0; i < nr; ++i) h[(int64_t)i * nc + k] = src[i];
    }
    int beta_mode = 0;
    double beta_scalar = 0.0;
    if (beta != nullptr) {
        if (incb == 0) {
            beta_mode = 1;
            beta_scalar = *beta;
        } else {
            beta_mode = 2;
            for (int k = 0; k < nc; ++k) h[ncoef + k] = beta[(int64_t)k * incb];
        }
    }
    double *d = g.d;
    HIPCHECK(hipMemcpyAsync(d, h, sizeof(double) * (ncoef + (beta_mode == 2 ? nc : 0)),
                            hipMemcpyHostToDevice, c.stream));
    lincomb_launch(c, X->d + (int64_t)start[0] * X->ldd, X->ldd, Y->nrows, nr, nc, d, library,
                   beta_mode, beta_scalar, d + ncoef, ys, Y->ldd);
    // no wait: the staging slot is reused only after this launch completes
    c.staged(g);
}

// One pass of block Cholesky QR of x[:, sx:sx+n] in the B inner product
// (B null: the plain one) with its small dense step on the device: B x into
// mv_ws[:, 0:n], the Gram G = x^T B x (lower triangle), the coefficients
// C = D^{-1} R^{-1} from G (cholqr_coef), x C into mv_ws and back into x.
// The sequence of the host pass -- MultiVecQtAP('S', 'S') with B,
// Cholesky and triangular inverse on the host, MultiVecLinearComb and
// MultiVecAxpby(1, ws, 0, x) -- without the Gram's copy to the host; one
// synchronization reads the factorization's flag at the end.  Returns 1
// when done, 0 when the block is numerically dependent (x as it was), -1
// when n is outside 1..1024 (nothing done).
int cholqr_pass(void **x, int sx, int n, void *matB, void **mv_ws, double zero_tol,
                double rel_tol)
{
    if (n < 1 || n > 1024) return -1;
    MV *X = mv_of(x, "cholqr x"), *W = mv_of(mv_ws, "cholqr ws");
    check_range(X, sx, sx + n, "cholqr x");
    check_range(W, 0, n, "cholqr ws");
    if (W->nrows != X->nrows) fail("cholqr: row counts differ", __FILE__, __LINE__);
    Ctx &c = ctx();
    double *xs = X->d + (int64_t)sx * X->ldd;
    const double *bx = xs;
    int64_t ldbx = X->ldd;
    if (matB != nullptr) {
        int s[2] = {sx, 0}, e[2] = {sx + n, n};
        h_MatDotMultiVec(matB, x, mv_ws, s, e, nullptr);
        bx = W->d;
        ldbx = W->ldd;
    }
    // G, C, the flag and the factorization's work array in one buffer
    const int64_t nn = (int64_t)n * n;
    double *G = c.small(2 * nn + 64 + cholqr_coef_work(n));
    double *C = G + nn, *flag = G + 2 * nn, *work = G + 2 * nn + 64;
    double *h = c.pinned(1);
    const GramPlan plan = gram_plan(X->nrows, n, n, true);
    double *part = c.scratch(plan.scratch);
    hipEvent_t ev = c.prof.begin(PROF_GRAM, 8.0 * (double)X->nrows * 2.0 * n,
                                 2.0 * (double)X->nrows * n * n, c.stream);
    gram(plan, X->nrows, n, n, xs, X->ldd, bx, ldbx, G, n, true, part, c.stream);
    c.prof.end(ev, c.stream);
    c.stats.gram_calls++;
    c.stats.gram_flops += (double)X->nrows * n * n;
    const bool library = lincomb_library(X, W, n, n, false);
    if (!cholqr_coef(n, G, n, zero_tol, rel_tol, work, C, library, flag, c.stream))
        fail("cholqr: order not handled", __FILE__, __LINE__);
    lincomb_launch(c, xs, X->ldd, X->nrows, n, n, C, library, 0, 0.0, nullptr, W->d, W->ldd);
    axpby(X->nrows, n, 1.0, W->d, W->ldd, 0.0, xs, X->ldd, c.stream);
    HIPCHECK(hipMemcpyAsync(h, flag, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return h[0] == 0.0 ? 1 : 0;
}

// The P step of GCG (src/ops_eig_sol_gcg.c ComputeP) with its small dense
// work on the device: the coefficient columns Z[:, sp:sp+np] (host, n rows,
// ld n) are orthonormalized against Z[:, 0:sp] and among themselves in the
// identity inner product -- the earlier columns removed twice, then
// Cholesky QR twice with one more removal between the passes, the scheme
// of the block orthonormalization of W -- and written back into Z, and
// ws[:, 0:np] = V[:, sv:sv+n] times them is copied into V[:, dst:dst+np].
// One synchronization reads both passes' flags and the columns before the
// product.  Returns 1 when done, 0 when a pass finds the columns
// numerically dependent (V, ws and Z untouched: the caller orthonormalizes
// Z on the host), -1 when ops is not this table or np is outside 1..1024
// (nothing done).
int orth_coef_comb_slot(OPS *ops, double *Z, int n, int sp, int np, double zero_tol,
                        void **v, int sv, void **ws, int dst)
{
    if (ops == nullptr || ops->MatDotMultiVec != h_MatDotMultiVec || np < 1 || np > 1024 ||
        n < 1 || sp < 0)
        return -1;
    MV *V = mv_of(v, "ComputeP V"), *W = mv_of(ws, "ComputeP ws");
    check_range(V, sv, sv + n, "ComputeP V");
    check_range(V, dst, dst + np, "ComputeP V");
    check_range(W, 0, np, "ComputeP ws");
    if (W->nrows != V->nrows) fail("ComputeP: row counts differ", __FILE__, __LINE__);
    Ctx &c = ctx();
    // device copy of Z[:, 0:sp+np]; the staging slot stays in use until the
    // event recorded after the last launch that reads it
    const int64_t nz = (int64_t)n * (sp + np);
    Ctx::Staging &g = c.staging(nz);
    memcpy(g.h, Z, sizeof(double) * nz);
    HIPCHECK(hipMemcpyAsync(g.d, g.h, sizeof(double) * nz, hipMemcpyHostToDevice, c.stream));
    const double *Zx = g.d;
    double *Zp = g.d + (int64_t)n * sp;
    // T (n x np), Cx (sp x np), G and C (np x np), the coefficient
    // transpose (np x n), the flags and the factorization's work array
    const int64_t nT = (int64_t)n * np, nC = (int64_t)np * np;
    double *T = c.small(2 * nT + (int64_t)sp * np + 2 * nC + 64 + cholqr_coef_work(np));
    double *Ft = T + nT, *Cx = Ft + nT, *G = Cx + (int64_t)sp * np, *C = G + nC,
           *flag = C + nC, *work = flag + 64;
    double *h = c.pinned(2);
    auto remove_previous = [&](double *Y) {
        if (sp == 0) return;
        gemm_small(true, sp, np, n, 1.0, Zx, n, Y, n, 0.0, Cx, sp);
        gemm_small(false, n, np, sp, -1.0, Zx, n, Cx, sp, 1.0, Y, n);
    };
    auto pass = [&](const double *Y, double *out, double *f) {
        gemm_small(true, np, np, n, 1.0, Y, n, Y, n, 0.0, G, np);
        if (!cholqr_coef(np, G, np, zero_tol, 1e-10, work, C, true, f, c.stream))
            fail("ComputeP: order not handled", __FILE__, __LINE__);
        gemm_small(false, n, np, np, 1.0, Y, n, C, np, 0.0, out, n);
    };
    remove_previous(Zp);
    remove_previous(Zp);
    pass(Zp, T, flag);
    remove_previous(T);
    pass(T, Zp, flag + 1);
    // the orthonormalized columns go back to Z too: the next Rayleigh-Ritz
    // step projects with them on the host (P^T A P from the coefficients)
    HIPCHECK(hipMemcpyAsync(g.h + (int64_t)n * sp, Zp, sizeof(double) * nT,
                            hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipMemcpyAsync(h, flag, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (h[0] != 0.0 || h[1] != 0.0) {
        c.staged(g);
        return 0;
    }
    memcpy(Z + (int64_t)n * sp, g.h + (int64_t)n * sp, sizeof(double) * nT);
    const bool library = lincomb_library(V, W, n, np, false);
    const double *coef = Zp;
    if (!library) {
        transpose(n, np, Zp, n, Ft, np);
        coef = Ft;
    }
    lincomb_launch(c, V->d + (int64_t)sv * V->ldd, V->ldd, V->nrows, n, np, coef, library, 0, 0.0,
                   nullptr, W->d, W->ldd);
    axpby(V->nrows, np, 1.0, W->d, W->ldd, 0.0, V->d + (int64_t)dst * V->ldd, V->ldd, c.stream);
    c.staged(g);
    return 1;
}

int cholqr_pass_slot(OPS *ops, void **x, int sx, int n, void *matB, void **mv_ws,
                     double zero_tol, double rel_tol)
{
    if (ops == nullptr || ops->MatDotMultiVec != h_MatDotMultiVec) return -1;
    return cholqr_pass(x, sx, n, matB, mv_ws, zero_tol, rel_tol);
}

// Inner products to a host array (app/app_lapack.c:299-321, 24-183)
void inner_prod_impl(char nsd, MV *X, MV *Y, int *start, int *end, double *ip, int ldIP)
{
    const int nr = end[0] - start[0], nc = end[1] - start[1];
    if (nr <= 0 || nc <= 0) return;
    check_range(X, start[0], end[0], "MultiVecInnerProd x");
    check_range(Y, start[1], end[1], "MultiVecInnerProd y");
    if (X->nrows != Y->nrows) fail("MultiVecInnerProd: row counts differ", __FILE__, __LINE__);
    Ctx &c = ctx();
    const double *xs = X->d + (int64_t)start[0] * X->ldd;
    const double *ys = Y->d + (int64_t)start[1] * Y->ldd;
    if (nsd == 'D') {
        if (nr != nc) fail("MultiVecInnerProd 'D': ranges differ", __FILE__, __LINE__);
        double *part = c.scratch(dots_scratch(nc));
        double *d = c.small(nc);
        col_dots(X->nrows, nc, xs, X->ldd, ys, Y->ldd, d, part, c.stream);
        double *h = c.pinned(nc);
        HIPCHECK(hipMemcpyAsync(h, d, sizeof(double) * nc, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        for (int k = 0; k < nc; ++k) ip[(int64_t)ldIP * k] = h[k];
        c.stats.dot_calls++;
        return;
    }
    const bool sym = (nsd == 'S');
    if (sym && nr != nc) fail("MultiVecInnerProd 'S': block not square", __FILE__, __LINE__);
    const GramPlan plan = gram_plan(X->nrows, nr, nc, sym);
    double *part = c.scratch(plan.scratch);
    double *d = c.small((int64_t)nr * nc);
    hipEvent_t ev = c.prof.begin(PROF_GRAM, 8.0 * (double)X->nrows * (nr + nc),
                                 2.0 * (double)X->nrows * nr * nc, c.stream);
    gram(plan, X->nrows, nr, nc, xs, X->ldd, ys, Y->ldd, d, nr, sym, part, c.stream);
    c.prof.end(ev, c.stream);
    double *h = c.pinned((int64_t)nr * nc);
    HIPCHECK(hipMemcpyAsync(h, d, sizeof(double) * nr * nc, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    c.stats.gram_calls++;
    c.stats.gram_flops += 2.0 * (double)X->nrows * nr * nc * (sym ? 0.5 : 1.0);
    for (int j = 0; j < nc; ++j) {
        if (sym) {
            for (int i = j; i < nr; ++i) {
                const double v = h[i + (int64_t)nr * j];
                ip[i + (int64_t)ldIP * j] = v;
                ip[j + (int64_t)ldIP * i] = v;
            }
        } else {
            for (int i = 0; i < nr; ++i) ip[i + (int64_t)ldIP * j] = h[i + (int64_t)nr * j];
        }
    }
}

void h_MultiVecInnerProd(char nsdIP, void **x, void **y, int, int *start, int *end,
                         double *ip, int ldIP, OPS *)
{
    inner_prod_impl(nsdIP, mv_of(x, "MultiVecInnerProd x"), mv_of(y, "MultiVecInnerProd y"),
                    start, end, ip, ldIP);
}

// qAp = Q^T op(A) P with op(A) P left in mv_ws (src/ops_multi_vec.c:351-411)
void h_MultiVecQtAP(char ntsA, char ntsdQAP, void **mvQ, void *matA, void **mvP, int is_vec,
                    int *startQP, int *endQP, double *qAp, int ldQAP, void **mv_ws, OPS *ops)
{
    const int nr = endQP[0] - startQP[0], nc = endQP[1] - startQP[1];
    if (nr <= 0 || nc <= 0) return;
    int s[2], e[2];
    if (matA == nullptr) {
        if (ntsdQAP == 'T') {
            s[0] = startQP[1]; e[0] = endQP[1];
            s[1] = startQP[0]; e[1] = endQP[0];
            h_MultiVecInnerProd('N', mvP, mvQ, is_vec, s, e, qAp, ldQAP, ops);
        } else {
            h_MultiVecInnerProd(ntsdQAP, mvQ, mvP, is_vec, startQP, endQP, qAp, ldQAP, ops);
        }
        return;
    }
    // op(A) = A: the backend requires symmetric matrices ('T' == 'N')
    (void)ntsA;
    s[0] = startQP[1]; e[0] = endQP[1];
    s[1] = 0; e[1] = nc;
    h_MatDotMultiVec(matA, mvP, mv_ws, s, e, ops);
    if (ntsdQAP == 'T') {
        s[0] = 0; e[0] = nc;
        s[1] = startQP[0]; e[1] = endQP[0];
        h_MultiVecInnerProd('N', mv_ws, mvQ, is_vec, s, e, qAp, ldQAP, ops);
    } else {
        s[0] = startQP[0]; e[0] = endQP[0];
        s[1] = 0; e[1] = nc;
        h_MultiVecInnerProd(ntsdQAP, mvQ, mv_ws, is_vec, s, e, qAp, ldQAP, ops);
    }
}

void h_MultiVecSetRandomValue(void **x, int start, int end, OPS *)
{
    MV *m = mv_of(x, "MultiVecSetRandomValue");
    check_range(m, start, end, "MultiVecSetRandomValue");
    const int nc = end - start;
    if (nc <= 0 || m->nrows == 0) return;
    Ctx &c = ctx();
    if (c.rand_mode == GCGE_HIP_RAND_DEVICE) {
        random_uniform(m->nrows, nc, m->d + (int64_t)start * m->ldd, m->ldd, c.rand_seed,
                       c.rand_counter, c.stream);
        c.rand_counter += (uint64_t)m->nrows * (uint64_t)nc;
        return;
    }
    // libc rand(), column by column (app/app_lapack.c:322-333)
    const int64_t n = m->nrows;
    double *h = c.pinned(n);
    for (int k = 0; k < nc; ++k) {
        for (int64_t r = 0; r < n; ++r) h[r] = (double)rand() / ((double)RAND_MAX + 1.0);
        HIPCHECK(hipMemcpyAsync(m->d + (int64_t)(start + k) * m->ldd, h, sizeof(double) * n,
                                hipMemcpyHostToDevice, c.stream));
        c.sync();
    }
}

void h_MatView(void *mat, OPS *ops)
{
    Mat *A = mat_of(mat, "MatView");
    ops->Printf("HIP SELL-64 matrix: %lld x %lld, nnz %lld (stored %lld)\n", (long long)A->n,
                (long long)A->n, (long long)A->sell.nnz_true, (long long)A->sell.nnz);
}

// single vectors are one-column multivectors
void h_VecCreateByMat(void **v, void *m, OPS *ops)
{
    h_MultiVecCreateByMat(reinterpret_cast<void ***>(v), 1, m, ops);
}
void h_VecCreateByVec(void **v, void *s, OPS *ops)
{
    h_MultiVecCreateByVec(reinterpret_cast<void ***>(v), 1, s, ops);
}
void h_VecDestroy(void **v, OPS *ops) { h_MultiVecDestroy(reinterpret_cast<void ***>(v), 1, ops); }
void h_VecView(void *x, OPS *ops) { h_MultiVecView(reinterpret_cast<void **>(x), 0, 1, ops); }
void h_VecInnerProd(void *x, void *y, double *ip, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    h_MultiVecInnerProd('S', reinterpret_cast<void **>(x), reinterpret_cast<void **>(y), 0, s,
                        e, ip, 1, ops);
}
void h_VecSetRandomValue(void *x, OPS *ops)
{
    h_MultiVecSetRandomValue(reinterpret_cast<void **>(x), 0, 1, ops);
}
void h_VecAxpby(double a, void *x, double b, void *y, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    h_MultiVecAxpby(a, reinterpret_cast<void **>(x), b, reinterpret_cast<void **>(y), s, e, ops);
}
void h_MatDotVec(void *m, void *x, void *y, OPS *ops)
{
    int s[2] = {0, 0}, e[2] = {1, 1};
    h_MatDotMultiVec(m, reinterpret_cast<void **>(x), reinterpret_cast<void **>(y), s, e, ops);
}

}  // namespace

namespace gcge {
bool hip_stencil_residual(OPS *ops, void *A, void *B, void **x, int sx, int nc,
                          const double *lam, void **r, int sr)
{
    return stencil_residual_slot(ops, A, B, x, sx, nc, lam, r, sr);
}
int hip_cholqr_pass(OPS *ops, void **x, int sx, int n, void *B, void **mv_ws, double zero_tol,
                    double rel_tol)
{
    return cholqr_pass_slot(ops, x, sx, n, B, mv_ws, zero_tol, rel_tol);
}
int hip_orth_coef_comb(OPS *ops, double *Z, int n, int sp, int np, double zero_tol,
                       void **V, int sv, void **ws, int dst)
{
    return orth_coef_comb_slot(ops, Z, n, sp, np, zero_tol, V, sv, ws, dst);
}
}  // namespace gcge

// ------------------------------------------------------------------ C ABI
extern "C" {

void OPS_HIP_Clear(OPS *ops)
{
    if (ops->lapack_ops) {
        free(ops->lapack_ops);
        ops->lapack_ops = nullptr;
    }
}

void OPS_HIP_Set(OPS *ops)
{
    if (ops->lapack_ops != nullptr)
        fail("OPS_HIP_Set: ops->lapack_ops must be NULL (as for OPS_CCS_Set)", __FILE__,
             __LINE__);
    (void)ctx();
    OPS *lap = static_cast<OPS *>(calloc(1, sizeof(OPS)));
    OPS_HOSTDENSE_Set(lap);
    ops->lapack_ops = lap;
    ops->Printf = gcge_host_Printf;
    ops->GetWtime = gcge_host_GetWtime;
    ops->GetOptionFromCommandLine = gcge_host_GetOptionFromCommandLine;
    ops->MatView = h_MatView;
    ops->MatAxpby = nullptr;
    ops->VecCreateByMat = h_VecCreateByMat;
    ops->VecCreateByVec = h_VecCreateByVec;
    ops->VecDestroy = h_VecDestroy;
    ops->VecView = h_VecView;
    ops->VecInnerProd = h_VecInnerProd;
    ops->VecLocalInnerProd = h_VecInnerProd;
    ops->VecSetRandomValue = h_VecSetRandomValue;
    ops->VecAxpby = h_VecAxpby;
    ops->MatDotVec = h_MatDotVec;
    ops->MatTransDotVec = h_MatDotVec;
    ops->MultiVecCreateByMat = h_MultiVecCreateByMat;
    ops->MultiVecCreateByVec = h_MultiVecCreateByVec;
    ops->MultiVecCreateByMultiVec = h_MultiVecCreateByMultiVec;
    ops->MultiVecDestroy = h_MultiVecDestroy;
    ops->GetVecFromMultiVec = h_GetVecFromMultiVec;
    ops->RestoreVecForMultiVec = h_RestoreVecForMultiVec;
    ops->MultiVecView = h_MultiVecView;
    ops->MultiVecLocalInnerProd = h_MultiVecInnerProd;
    ops->MultiVecInnerProd = h_MultiVecInnerProd;
    ops->MultiVecSetRandomValue = h_MultiVecSetRandomValue;
    ops->MultiVecAxpby = h_MultiVecAxpby;
    ops->MultiVecLinearComb = h_MultiVecLinearComb;
    ops->MatDotMultiVec = h_MatDotMultiVec;
    ops->MatTransDotMultiVec = h_MatDotMultiVec;
    ops->MultiVecQtAP = h_MultiVecQtAP;
    ops->DenseMatQtAP = lap->DenseMatQtAP;
    ops->DenseMatOrth = lap->DenseMatOrth;
}

// The grid side n when the CCS matrix is a constant-coefficient 7-point
// operator on an n^3 Dirichlet grid (every row: its in-grid neighbours
// x +- 1, y +- n, z +- n^2 and itself in ascending order, one value on the
// diagonal and one on every neighbour), with those values; 0 otherwise.
static int detect_stencil7(const gcge_ccs_mat *m, double *d, double *o)
{
    const int64_t N = m->nrows;
    int64_t n = (int64_t)std::llround(std::cbrt((double)N));
    if (n < 4 || n > kStencilMaxN || n * n * n != N) return 0;
    const int *jc = m->j_col, *ir = m->i_row;
    const double *v = m->data;
    double dv = 0.0, ov = 0.0;
    bool have_o = false;
    for (int64_t r = 0; r < N; ++r) {
        const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
        int64_t want[7];
        int k = 0;
        if (z > 0) want[k++] = r - n * n;
        if (y > 0) want[k++] = r - n;
        if (x > 0) want[k++] = r - 1;
        want[k++] = r;
        if (x < n - 1) want[k++] = r + 1;
        if (y < n - 1) want[k++] = r + n;
        if (z < n - 1) want[k++] = r + n * n;
        if (jc[r + 1] - jc[r] != k) return 0;
        for (int j = 0; j < k; ++j) {
            const int64_t e = jc[r] + j;
            if (ir[e] != want[j]) return 0;
            const double val = v[e];
            if (want[j] == r) {
                if (r == 0) dv = val;
                else if (val != dv || std::signbit(val) != std::signbit(dv)) return 0;
            } else {
                if (!have_o) {
                    ov = val;
                    have_o = true;
                } else if (val != ov || std::signbit(val) != std::signbit(ov)) {
                    return 0;
                }
            }
        }
    }
    *d = dv;
    *o = ov;
    return (int)n;
}

void *HIPCCS_Create(const gcge_ccs_mat *host, int flags)
{
    if (host == nullptr || host->nrows <= 0 || host->nrows != host->ncols ||
        host->j_col == nullptr || host->i_row == nullptr || host->data == nullptr)
        return nullptr;
    const int64_t n = host->nrows;
    const int *jc = host->j_col;
    if (jc[0] != 0) return nullptr;
    for (int64_t j = 0; j < n; ++j)
        if (jc[j + 1] < jc[j]) return nullptr;
    const int64_t nnz = jc[n];
    for (int64_t e = 0; e < nnz; ++e)
        if (host->i_row[e] < 0 || host->i_row[e] >= n) return nullptr;
    Ctx &c = ctx();
    Mat *M = new Mat();
    M->n = n;
    SellMat &S = M->sell;
    S.nrows = n;
    S.nslices = (n + 63) / 64;
    S.nnz_true = nnz;
    std::vector<int64_t> sp((size_t)S.nslices + 1, 0);
    int maxw = 0;
    for (int64_t s = 0; s < S.nslices; ++s) {
        int w = 0;
        for (int64_t r = s * 64; r < (s + 1) * 64 && r < n; ++r) {
            const int len = jc[r + 1] - jc[r];
            if (len > w) w = len;
        }
        if (w > maxw) maxw = w;
        sp[s + 1] = sp[s] + (int64_t)w * 64;
    }
    S.max_width = maxw;
    S.nnz = sp[S.nslices];
    S.structure_hash = slice_hash(sp);
    // Column r of a symmetric CCS matrix is row r: fill SELL by rows.
    std::vector<int> colv((size_t)(S.nnz > 0 ? S.nnz : 1));
    std::vector<double> valv((size_t)(S.nnz > 0 ? S.nnz : 1));
    for (int64_t s = 0; s < S.nslices; ++s) {
        const int w = (int)((sp[s + 1] - sp[s]) / 64);
        for (int l = 0; l < 64; ++l) {
            const int64_t r = s * 64 + l;
            const int len = r < n ? jc[r + 1] - jc[r] : 0;
            for (int j = 0; j < w; ++j) {
                const int64_t e = sp[s] + (int64_t)j * 64 + l;
                if (j < len) {
                    colv[e] = host->i_row[jc[r] + j];
                    valv[e] = host->data[jc[r] + j];
                } else {
                    colv[e] = (int)(r < n ? r : 0);
                    valv[e] = 0.0;
                }
            }
        }
    }
    dev_alloc(&S.slice_ptr, sizeof(int64_t) * sp.size());
    dev_alloc(&S.col, sizeof(int) * colv.size());
    dev_alloc(&S.val, sizeof(double) * valv.size());
    HIPCHECK(hipMemcpy(S.slice_ptr, sp.data(), sizeof(int64_t) * sp.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(S.col, colv.data(), sizeof(int) * colv.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(S.val, valv.data(), sizeof(double) * valv.size(), hipMemcpyHostToDevice));
    M->bytes = (int64_t)(sizeof(int64_t) * sp.size() + sizeof(int) * colv.size() +
                         sizeof(double) * valv.size());
    if (flags & GCGE_HIP_MAT_FP32) {
        std::vector<float> v32(valv.size());
        for (size_t i = 0; i < valv.size(); ++i) v32[i] = (float)valv[i];
        dev_alloc(&S.val32, sizeof(float) * v32.size());
        HIPCHECK(hipMemcpy(S.val32, v32.data(), sizeof(float) * v32.size(), hipMemcpyHostToDevice));
        M->bytes += (int64_t)(sizeof(float) * v32.size());
    }
    detect_plane(S, c.stream);
    if (!(flags & GCGE_HIP_MAT_GENERAL))
        S.stencil_n = detect_stencil7(host, &S.stencil_d, &S.stencil_o);
    return M;
}

void HIPCCS_Destroy(void **mat)
{
    if (mat == nullptr || *mat == nullptr) return;
    Mat *M = mat_of(*mat, "HIPCCS_Destroy");
    ctx().sync();
    dev_free(M->sell.slice_ptr);
    dev_free(M->sell.col);
    dev_free(M->sell.val);
    if (M->sell.val32) dev_free(M->sell.val32);
    M->magic = 0;
    delete M;
    *mat = nullptr;
}

int HIPCCS_Stencil(void *mat) { return mat_of(mat, "HIPCCS_Stencil")->sell.stencil_n; }

void HIPCCS_Info(void *mat, int64_t *nrows, int64_t *nnz, int64_t *padded, int64_t *bytes)
{
    Mat *M = mat_of(mat, "HIPCCS_Info");
    if (nrows) *nrows = M->n;
    if (nnz) *nnz = M->sell.nnz_true;
    if (padded) *padded = M->sell.nnz;
    if (bytes) *bytes = M->bytes;
}

void HIP_MultiVecSetFromHost(void **mv, int start, int end, const double *src, int64_t ld)
{
    MV *m = mv_of(mv, "HIP_MultiVecSetFromHost");
    check_range(m, start, end, "HIP_MultiVecSetFromHost");
    Ctx &c = ctx();
    for (int k = 0; k < end - start; ++k)
        HIPCHECK(hipMemcpyAsync(m->d + (int64_t)(start + k) * m->ldd, src + (int64_t)k * ld,
                                sizeof(double) * m->nrows, hipMemcpyHostToDevice, c.stream));
    c.sync();
}

void HIP_MultiVecGetToHost(void **mv, int start, int end, double *dst, int64_t ld)
{
    MV *m = mv_of(mv, "HIP_MultiVecGetToHost");
    check_range(m, start, end, "HIP_MultiVecGetToHost");
    Ctx &c = ctx();
    for (int k = 0; k < end - start; ++k)
        HIPCHECK(hipMemcpyAsync(dst + (int64_t)k * ld, m->d + (int64_t)(start + k) * m->ldd,
                                sizeof(double) * m->nrows, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

void HIP_MultiVecInfo(void **mv, int64_t *nrows, int64_t *ncols, int64_t *ldd)
{
    MV *m = mv_of(mv, "HIP_MultiVecInfo");
    if (nrows) *nrows = m->nrows;
    if (ncols) *ncols = m->ncols;
    if (ldd) *ldd = m->ldd;
}

void gcge_hip_set_random_mode(int mode, uint64_t seed)
{
    Ctx &c = ctx();
    c.rand_mode = mode;
    c.rand_seed = seed;
    c.rand_counter = 0;
}

int HIP_MultiVecCholQR(void **x, int start, int end, void *B, void **mv_ws, double zero_tol,
                       double rel_tol)
{
    return cholqr_pass(x, start, end - start, B, mv_ws, zero_tol, rel_tol);
}

void gcge_hip_synchronize(void) { ctx().sync(); }

void gcge_hip_thread_release(void) { release_ctx(); }

int gcge_hip_last_error(char *buf, int len)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    if (buf && len > 0) {
        strncpy(buf, g_err.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return g_err_code;
}

int gcge_hip_device(void) { return ctx().device; }

void gcge_hip_set_memory_pool(int on)
{
    Ctx &c = ctx();
    if (on) {
        if (c.pool == nullptr) c.pool = new MemPool();
        std::lock_guard<std::mutex> lk(g_blk_mu);
        c.pool->on = true;
    } else if (c.pool != nullptr) {
        c.sync();
        pool_flush(c.pool, false);
    }
}

void gcge_hip_memory_pool_stats(int64_t *allocated, int64_t *cached)
{
    Ctx &c = ctx();
    int64_t a = 0;
    std::lock_guard<std::mutex> lk(g_blk_mu);
    for (auto &e : g_blocks)
        if (e.second.pool == c.pool && c.pool != nullptr) a += (int64_t)e.second.size;
    if (allocated) *allocated = a;
    if (cached) *cached = c.pool ? (int64_t)c.pool->cached_bytes : 0;
}

void gcge_hip_profile_enable(int on)
{
    Ctx &c = ctx();
    c.prof.collect();
    c.prof.on = on != 0;
}

void gcge_hip_profile_reset(void) { ctx().prof.reset(); }

void gcge_hip_profile_all(int on)
{
    std::lock_guard<std::mutex> lp(g_prof_mu);
    g_prof_all = on != 0;
    if (on)
        for (int k = 0; k < PROF_NKIND; ++k) {
            g_prof_launches[k] = 0;
            g_prof_ms[k] = g_prof_bytes[k] = g_prof_flops[k] = 0.0;
        }
}

void gcge_hip_profile_read_all(int kind, int64_t *launches, double *total_ms, double *bytes,
                               double *flops)
{
    if (kind < 0 || kind >= PROF_NKIND)
        fail("gcge_hip_profile_read_all: bad kind", __FILE__, __LINE__);
    std::lock_guard<std::mutex> lp(g_prof_mu);
    if (launches) *launches = g_prof_launches[kind];
    if (total_ms) *total_ms = g_prof_ms[kind];
    if (bytes) *bytes = g_prof_bytes[kind];
    if (flops) *flops = g_prof_flops[kind];
}

void gcge_hip_profile_read(int kind, int64_t *launches, double *total_ms, double *bytes,
                           double *flops)
{
    Ctx &c = ctx();
    c.prof.collect();
    if (kind < 0 || kind >= PROF_NKIND) fail("gcge_hip_profile_read: bad kind", __FILE__, __LINE__);
    if (launches) *launches = c.prof.launches[kind];
    if (total_ms) *total_ms = c.prof.ms[kind];
    if (bytes) *bytes = c.prof.bytes[kind];
    if (flops) *flops = c.prof.flops[kind];
}

}  // extern "C"

// Solve service: generalized eigenproblems served concurrently on one GPU.
//
// Requests (a CCS pair and the GCG sizes) enter a FIFO queue.  A pool of
// worker threads, each with its own device stream (the library's contexts
// are per thread), takes the request at the head of the queue as soon as a
// worker is free and the request's device footprint fits in what the HBM
// budget has left, so requests that arrive while others are on the GPU
// are admitted as slots and memory free up and their work overlaps on the
// device.  Each request runs this library's GCG driver over its own
// operation table, FP64 or mixed precision, with the initial block from
// the device generator keyed by the request's seed: a result does not
// depend on what else runs.
//
// Nothing here has a counterpart in the reference, which solves one
// problem per process (gcge.c:442-535); the per-request call sequence is
// the one of that plug-in.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "gcge_hip.h"

namespace gcge {
// The batched BlockPCG of pcg32.hip: one thread running the inner
// iterations of every mixed-precision solve attached to it in shared
// launches.
void *pcg32_batcher_create();
void pcg32_batcher_destroy(void *b);
void pcg32_batcher_stats(void *b, int64_t *rounds, int64_t *job_iters, int *max_jobs);
void pcg32_set_batcher(OPS *ops, void *b);
}  // namespace gcge

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    gcge_request req;
    gcge_result *res = nullptr;
    int64_t footprint = 0;
    double t_submit = 0.0;
    bool done = false;
};

struct Service {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::map<int64_t, Job *> jobs;
    int64_t budget = 0, in_use = 0, next_ticket = 0;
    int running = 0, max_running = 0;
    int64_t completed = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    // the inner iterations of mixed requests run in one batcher's shared
    // launches while batching is on (the default)
    void *batcher = nullptr;
    bool batching = true;
    // block orthonormalization of the mixed requests: Cholesky QR with its
    // factorizations on the host, which served the config-4 and config-5
    // streams 2-6 % faster than on the device (16 workers: a device pass
    // waits for its product before the host can go on, a host pass only
    // for its Gram, and the host factorization overlaps other requests'
    // kernels)
    int orth_mixed = GCGE_ORTH_CHOLQR_HOST;
};

int64_t nnz_of(const gcge_ccs_mat *m) { return m ? (int64_t)m->j_col[m->ncols] : 0; }

// One request on the calling thread's stream.  Returns 0 or a negative
// status (matrix rejected).
int run_one(Service *s, Job &j)
{
    const gcge_request &q = j.req;
    gcge_result &r = *j.res;
    const bool mixed = q.mode == GCGE_MODE_MIXED;
    const int flags = mixed ? GCGE_HIP_MAT_FP32 : 0;
    if (q.nev_given < 0 || q.nev_given > q.nev_init || (q.nev_given > 0 && q.x0 == nullptr))
        return -1;
    const double t0 = now_s();
    void *A = HIPCCS_Create(q.A, flags);
    if (A == nullptr) return -1;
    void *B = nullptr;
    if (q.B != nullptr) {
        B = HIPCCS_Create(q.B, flags);
        if (B == nullptr) {
            HIPCCS_Destroy(&A);
            return -1;
        }
    }
    OPS *ops = static_cast<OPS *>(calloc(1, sizeof(OPS)));
    OPS_HIP_Set(ops);
    void **evec = nullptr;
    ops->MultiVecCreateByMat(&evec, q.nev_max, A, ops);
    void **mv_ws[4] = {nullptr, nullptr, nullptr, nullptr};
    HIPGCG_Create(A, q.nev_max, q.block_size, q.nev_init, mv_ws, nullptr, nullptr, ops);
    gcge_hip_set_random_mode(GCGE_HIP_RAND_DEVICE, q.seed);
    double tol[2] = {q.tol_abs, q.tol_rel};
    HIPGCG_EigenSolverSetup(q.gap_min, q.nev_init, q.nev_max, q.block_size, tol, q.max_iter, 0,
                            0, mixed ? 1 : 0, mv_ws, nullptr, nullptr, ops);
    HIPGCG_Setparameters(q.gap_min, mixed ? 0 : q.shift, ops);
    if (mixed) HIPGCG_MultiLinearSolverSetup_BlockPCG32(30, 1e-2, 1e-14, B, 0.0, ops);
    bool batched;
    int orth;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        batched = s->batching;
        orth = s->orth_mixed;
    }
    HIPGCG_SetOrthMethod(ops, mixed ? orth : GCGE_ORTH_MGS);
    if (mixed && batched) gcge::pcg32_set_batcher(ops, s->batcher);
    std::vector<double> eval((size_t)q.nev_max, 0.0);
    int nconv = q.nev;
    // warm start: the given vectors in the first columns of evec, where
    // the driver takes nevGiven initial vectors from
    if (q.nev_given > 0) HIP_MultiVecSetFromHost(evec, 0, q.nev_given, q.x0, q.A->nrows);
    gcge_hip_synchronize();
    const double t1 = now_s();
    ops->EigenSolver(A, B, eval.data(), evec, q.nev_given, &nconv, ops);
    gcge_hip_synchronize();
    const double t2 = now_s();
    int64_t stats[4];
    double times[10];
    HIPGCG_GetStats(ops, stats, times);
    r.nev_conv = nconv;
    r.num_iter = (int)stats[0];
    const int ncopy = nconv < q.nev_max ? nconv : q.nev_max;
    if (r.eval) memcpy(r.eval, eval.data(), sizeof(double) * (size_t)ncopy);
    if (r.evec && ncopy > 0) HIP_MultiVecGetToHost(evec, 0, ncopy, r.evec, q.A->nrows);
    HIPGCG_Destroy(mv_ws, nullptr, nullptr, q.nev_max, q.block_size, ops);
    ops->MultiVecDestroy(&evec, q.nev_max, ops);
    HIPGCG_Free(ops);
    OPS_HIP_Clear(ops);
    free(ops);
    HIPCCS_Destroy(&A);
    if (B) HIPCCS_Destroy(&B);
    gcge_hip_synchronize();
    gcge_hip_memory_pool_stats(&r.pool_bytes, nullptr);
    r.setup_seconds = t1 - t0;
    r.solve_seconds = t2 - t1;
    return 0;
}

void worker(Service *s)
{
    // device buffers freed by one request are kept for the next ones, so a
    // running request does not allocate or free device memory once the
    // worker's pool holds what its requests need (hipFree waits for the
    // whole device, which would stall every other worker)
    gcge_hip_set_memory_pool(1);
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return (s->stop && s->queue.empty()) ||
                       (!s->queue.empty() && s->queue.front()->footprint <= s->budget - s->in_use);
            });
            if (s->queue.empty()) break;  // stopping
            j = s->queue.front();
            s->queue.pop_front();
            s->in_use += j->footprint;
            s->running += 1;
            if (s->running > s->max_running) s->max_running = s->running;
            j->res->queue_seconds = now_s() - j->t_submit;
            j->res->concurrent = s->running;
        }
        const int st = run_one(s, *j);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->res->status = st;
            j->res->total_seconds = now_s() - j->t_submit;
            j->done = true;
            s->in_use -= j->footprint;
            s->running -= 1;
            s->completed += 1;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    gcge_hip_thread_release();
}

}  // namespace

extern "C" {

int64_t gcge_request_footprint(const gcge_request *q)
{
    const int64_t n = q->A->nrows;
    const int64_t bs = q->block_size, nev_max = q->nev_max;
    // multivectors: V (nevMax + 2 bs), three bs-blocks and evec (nevMax),
    // column-major FP64
    int64_t bytes = 8 * n * (2 * nev_max + 5 * bs);
    // matrices: FP64 values and int32 columns (plus slack for slice
    // padding), slice pointers
    for (const gcge_ccs_mat *m : {q->A, q->B}) {
        if (m == nullptr) continue;
        const int64_t nnz = nnz_of(m);
        bytes += (12 * nnz) * 9 / 8 + 8 * (n / 64 + 2);
        if (q->mode == GCGE_MODE_MIXED) bytes += (4 * nnz) * 9 / 8;
    }
    // mixed: FP32 BlockPCG blocks (correction, residual, direction, product)
    if (q->mode == GCGE_MODE_MIXED) bytes += 4 * 4 * n * ((bs + 7) / 8 * 8);
    // inner-product partials and projected matrices
    bytes += 8 * (int64_t)(nev_max + 2 * bs) * (nev_max + 2 * bs) * 4 + (16 << 20);
    return bytes;
}

void *gcge_service_create(int workers, int64_t hbm_budget_bytes)
{
    if (workers < 1 || hbm_budget_bytes <= 0) return nullptr;
    Service *s = new Service();
    s->budget = hbm_budget_bytes;
    s->batcher = gcge::pcg32_batcher_create();
    for (int i = 0; i < workers; ++i) s->workers.emplace_back(worker, s);
    return s;
}

void gcge_service_set_batching(void *svc, int on)
{
    Service *s = static_cast<Service *>(svc);
    std::lock_guard<std::mutex> lk(s->mu);
    s->batching = on != 0;
}

int gcge_service_set_orth(void *svc, int method)
{
    if (method != GCGE_ORTH_CHOLQR && method != GCGE_ORTH_CHOLQR_HOST) return -1;
    Service *s = static_cast<Service *>(svc);
    std::lock_guard<std::mutex> lk(s->mu);
    s->orth_mixed = method;
    return 0;
}

void gcge_service_batch_stats(void *svc, int64_t *rounds, int64_t *solve_iterations,
                              int *max_solves)
{
    Service *s = static_cast<Service *>(svc);
    gcge::pcg32_batcher_stats(s->batcher, rounds, solve_iterations, max_solves);
}

int64_t gcge_service_submit(void *svc, const gcge_request *req, gcge_result *res)
{
    Service *s = static_cast<Service *>(svc);
    if (s == nullptr || req == nullptr || res == nullptr || req->A == nullptr) return -1;
    Job *j = new Job();
    j->req = *req;
    j->res = res;
    j->footprint = gcge_request_footprint(req);
    memset(res, 0, offsetof(gcge_result, eval));
    res->status = 1;  // pending
    if (j->footprint > s->budget) {
        res->status = -2;  // can never be admitted
        delete j;
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        delete j;
        return -1;
    }
    j->t_submit = now_s();
    const int64_t t = s->next_ticket++;
    s->jobs[t] = j;
    s->queue.push_back(j);
    s->cv_work.notify_all();
    return t;
}

int gcge_service_wait(void *svc, int64_t ticket)
{
    Service *s = static_cast<Service *>(svc);
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return -3;
    Job *j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    const int st = j->res->status;
    s->jobs.erase(it);
    delete j;
    return st;
}

void gcge_service_stats(void *svc, int64_t *completed, int *max_concurrent, int64_t *hbm_in_use)
{
    Service *s = static_cast<Service *>(svc);
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (max_concurrent) *max_concurrent = s->max_running;
    if (hbm_in_use) *hbm_in_use = s->in_use;
}

void gcge_service_destroy(void *svc)
{
    Service *s = static_cast<Service *>(svc);
    if (s == nullptr) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    gcge::pcg32_batcher_destroy(s->batcher);  // after the workers: no solve is waiting
    for (auto &kv : s->jobs) delete kv.second;
    delete s;
}

}  // extern "C"

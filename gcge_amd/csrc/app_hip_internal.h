// Internal types of the HIP backend shared by the app, the driver and the
// mixed-precision solver.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "kernels.h"

namespace gcge {

[[noreturn]] void fail(const char *what, const char *file, int line);
// Records a failure for gcge_hip_last_error without aborting (entry points
// that report errors by return value).
void record_error(const char *what);
void hip_check(hipError_t e, const char *expr, const char *file, int line);
#define HIPCHECK(x) ::gcge::hip_check((x), #x, __FILE__, __LINE__)

// Device and pinned host memory of the library.  A thread may keep a pool
// (gcge_hip_set_memory_pool): while it is on, blocks the thread allocated
// are cached when freed, from any thread, and later allocations of the
// thread reuse a cached block of at least the size asked (and at most about
// twice it) before asking HIP, so a thread that runs request after request
// stops allocating once its pool holds what a request needs.  Without a
// pool these are hipMalloc / hipFree and hipHostMalloc / hipHostFree.
void *dev_alloc_bytes(size_t bytes);
void *pinned_alloc_bytes(size_t bytes);
void dev_free(void *p);
void pinned_free(void *p);
template <typename T>
void dev_alloc(T **p, size_t bytes)
{
    *p = static_cast<T *>(dev_alloc_bytes(bytes));
}
template <typename T>
void pinned_alloc(T **p, size_t bytes)
{
    *p = static_cast<T *>(pinned_alloc_bytes(bytes));
}
struct MemPool;

struct Stats {
    int64_t spmm_calls = 0, spmm_cols = 0, gram_calls = 0, lincomb_calls = 0, dot_calls = 0;
    double gram_flops = 0.0, lincomb_flops = 0.0;
};

// Per-kernel-class timing with HIP events recorded on the library stream
// around each launch; elapsed times are collected when read.
enum ProfKind { PROF_SPMM64 = 0, PROF_SPMM32 = 1, PROF_GRAM = 2, PROF_LINCOMB = 3,
                PROF_BLAS1 = 4, PROF_PCGVEC = 5, PROF_GEMM = 6, PROF_NKIND = 7 };
struct Profiler {
    bool on = false;
    struct Rec {
        hipEvent_t a, b;
        int kind;
        double bytes, flops;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    int64_t launches[PROF_NKIND] = {0};
    double ms[PROF_NKIND] = {0}, bytes[PROF_NKIND] = {0}, flops[PROF_NKIND] = {0};
    // returns the event to record after the launch, or null when off
    hipEvent_t begin(int kind, double bytes, double flops, hipStream_t s);
    void end(hipEvent_t e, hipStream_t s);
    void collect();
    void reset();
};

// One context per host thread: a stream and the thread's staging buffers.
// Every slot enqueues on the calling thread's stream and only slots that
// return host data synchronize, so tables driven from different threads
// run concurrently on the device.
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *d_scratch = nullptr;
    int64_t scratch_cap = 0;
    double *d_small = nullptr;
    int64_t small_cap = 0;
    double *h_pinned = nullptr;
    int64_t pinned_cap = 0;
    int rand_mode = 0;
    uint64_t rand_seed = 0;
    uint64_t rand_counter = 0;
    Stats stats;
    Profiler prof;
    // the device dense eigensolver's handle and buffers (dense_eig.cpp)
    void *dev_eig = nullptr;
    // the thread's memory pool, or null (dev_alloc_bytes)
    MemPool *pool = nullptr;
    // Upload staging for the coefficients of asynchronous slots: a ring of
    // pinned host / device buffer pairs, each reused only after the event
    // recorded behind the launches that read it has completed.
    struct Staging {
        double *h = nullptr, *d = nullptr;
        int64_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    static constexpr int kStage = 8;
    Staging stage[kStage];
    int stage_next = 0;
    Staging &staging(int64_t ndouble);
    void staged(Staging &s);
    double *scratch(int64_t ndouble);
    double *small(int64_t ndouble);
    double *pinned(int64_t ndouble);
    void sync();
};
Ctx &ctx();
// Frees the calling thread's context (a later call creates a new one).
void release_ctx();
// Stream capture is process-wide state in the HIP runtime: while one thread
// captures, allocations and synchronous calls made by other threads can be
// refused or can invalidate the capture.  The returned lock owns the
// context registry when the calling thread's context is the only one in
// the process, and is empty otherwise; a caller captures only while it
// owns the lock, and no other thread can create a context until it is
// released.
std::unique_lock<std::mutex> sole_context_lock();
void release_dev_eig(void *p);

// Device multivector: column-major nrows x ncols, leading dimension ldd.
struct MV {
    static constexpr uint32_t kMagic = 0x484d5631u;  // "HMV1"
    uint32_t magic = kMagic;
    int64_t nrows = 0, ncols = 0, ldd = 0;
    double *d = nullptr;
    bool owner = false;
};

// Device sparse matrix (SELL-64 of a symmetric CCS input).
struct Mat {
    static constexpr uint32_t kMagic = 0x484d5431u;  // "HMT1"
    uint32_t magic = kMagic;
    int64_t n = 0;
    int64_t bytes = 0;
    SellMat sell;
};

// Hash of a SELL matrix's slice pointers: equal hashes let A and B share
// one gather pass.
uint64_t slice_hash(const std::vector<int64_t> &sp);
MV *mv_of(void **p, const char *where);
Mat *mat_of(void *p, const char *where);
MV *mv_create(int64_t nrows, int64_t ncols);
void mv_destroy(MV *m);

}  // namespace gcge

// Serving loop: point clouds of uneven size arrive as requests; each runs
// knn_search(k) -> CSR -> compute_features on the device
// (pgeof_knn_features_device).
//
// - `n_slots` requests are in flight at once, each on its own HIP stream, so
//   their kernels and copies overlap on the GPU (continuous batching: a
//   request is admitted as soon as a slot and enough device memory are free,
//   not when a batch completes).
// - Device memory is one pool of hbm_budget_bytes, allocated when the server
//   is created.  An admitted request takes one contiguous range of it (first
//   fit) for its inputs, outputs and search workspace, and returns it when it
//   completes, so the server's device memory never exceeds the budget and
//   the workers never call hipMalloc or hipFree (hipFree would synchronise
//   the device and stall the other slots).
// - Admission is FIFO: the head request waits until a range of its
//   footprint is free; a request larger than the whole budget is rejected at
//   submit.  With nothing in flight the whole pool is free, so the head
//   request is always admitted eventually.
// - The grid and the CSR stay in device memory; the features (and the CSR
//   when the caller asks for it) are copied straight into the caller's
//   arrays.
// - Each request's result does not depend on what else is in flight: every
//   kernel is deterministic and a request's memory range is private to it.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pgeof_hip.h"

namespace
{

using Clock = std::chrono::steady_clock;

inline size_t al(size_t b) { return (b + 255) & ~size_t(255); }

// Device memory of one request: xyz, nn, nn_ptr, features, search workspace.
struct Carve
{
    size_t x, nn, ptr, f, ws, total;
};

Carve carve_of(int64_t n, uint32_t k)
{
    const size_t pts = (size_t)n;
    Carve c;
    c.x = al(pts * 12);
    c.nn = al(pts * k * 4);
    c.ptr = al((pts + 1) * 4);
    c.f = al(pts * 11 * 4);
    c.ws = al(pgeof_knn_workspace_size(n, n, k));
    c.total = c.x + c.nn + c.ptr + c.f + c.ws;
    return c;
}

struct Request
{
    uint64_t id;
    const float* xyz;
    int64_t n;
    uint32_t* nn_out;  // may be null
    float* feat_out;
    size_t footprint;
    size_t offset;  // start of the request's range in the pool (set at admission)
    Clock::time_point t_submit;
};

struct Done
{
    uint64_t id;
    int status;
    double latency_ms;
};

// First-fit allocator of ranges of the device pool (offsets only); callers
// hold the server mutex.
class RangePool
{
  public:
    void reset(size_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool take(size_t bytes, size_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes)
            {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void give(size_t off, size_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first)
        {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin())
        {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first)
            {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    bool fits(size_t bytes) const
    {
        for (const auto& kv : free_)
            if (kv.second >= bytes) return true;
        return false;
    }

  private:
    std::map<size_t, size_t> free_;  // offset -> size of each free range
};

}  // namespace

struct pgeof_server
{
    int device = 0;
    uint32_t k = 20;
    size_t budget = 0;
    char* pool_base = nullptr;
    RangePool pool;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_queue;  // queue / pool changes
    std::condition_variable cv_done;
    std::deque<Request> queue;
    std::deque<Done> done;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
};

namespace
{

int run_request(pgeof_server* S, hipStream_t st, const Request& r)
{
    const uint32_t k = S->k;
    const size_t n = (size_t)r.n;
    if (n == 0) return PGEOF_OK;
    const Carve c = carve_of(r.n, k);
    char* d = S->pool_base + r.offset;
    float* d_x = (float*)d;
    uint32_t* d_nn = (uint32_t*)(d + c.x);
    uint32_t* d_ptr = (uint32_t*)(d + c.x + c.nn);
    float* d_f = (float*)(d + c.x + c.nn + c.ptr);
    void* d_ws = d + c.x + c.nn + c.ptr + c.f;
    // straight from and into the caller's memory: the runtime stages pageable
    // memory itself, pipelined (measured faster than a pinned buffer per slot
    // plus host copies)
    int rc = PGEOF_OK;
    if (hipMemcpyAsync(d_x, r.xyz, n * 12, hipMemcpyHostToDevice, st) != hipSuccess) rc = PGEOF_ERUNTIME;
    if (!rc) rc = pgeof_knn_features_device(d_x, r.n, k, d_nn, d_ptr, nullptr, d_f, d_ws, c.ws, st);
    if (!rc && hipMemcpyAsync(r.feat_out, d_f, n * 11 * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = PGEOF_ERUNTIME;
    if (!rc && r.nn_out && hipMemcpyAsync(r.nn_out, d_nn, n * k * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = PGEOF_ERUNTIME;
    // every way out waits for the stream: the request is reported only once
    // nothing queued for it still reads the caller's arrays or its pool range
    const hipError_t se = hipStreamSynchronize(st);
    if (rc) return rc;
    return se == hipSuccess ? PGEOF_OK : PGEOF_ERUNTIME;
}

void worker(pgeof_server* S, int idx)
{
    (void)hipSetDevice(S->device);
    hipStream_t st = S->streams[idx];
    for (;;)
    {
        Request r;
        {
            std::unique_lock<std::mutex> l(S->mu);
            S->cv_queue.wait(l, [&] { return S->stop || (!S->queue.empty() && S->pool.fits(S->queue.front().footprint)); });
            if (S->stop) return;
            r = S->queue.front();
            S->queue.pop_front();
            if (!S->pool.take(r.footprint, &r.offset)) r.offset = 0;  // cannot happen: fits() was checked under the lock
        }
        S->cv_queue.notify_all();
        const int rc = run_request(S, st, r);
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - r.t_submit).count();
        {
            std::lock_guard<std::mutex> l(S->mu);
            S->pool.give(r.offset, r.footprint);
            S->done.push_back(Done{r.id, rc, ms});
        }
        S->cv_queue.notify_all();
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

pgeof_server* pgeof_server_create(int n_slots, size_t hbm_budget_bytes, uint32_t knn)
{
    if (n_slots < 1 || knn < 1 || hbm_budget_bytes == 0) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    pgeof_server* S = new pgeof_server();
    S->device = dev;
    S->k = knn;
    S->budget = hbm_budget_bytes & ~size_t(255);
    if (S->budget == 0 || hipMalloc(&S->pool_base, S->budget) != hipSuccess)
    {
        delete S;
        return nullptr;
    }
    S->pool.reset(S->budget);
    S->streams.assign((size_t)n_slots, nullptr);
    for (auto& s : S->streams)
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        {
            for (auto& t : S->streams)
                if (t) (void)hipStreamDestroy(t);
            (void)hipFree(S->pool_base);
            delete S;
            return nullptr;
        }
    for (int i = 0; i < n_slots; ++i) S->workers.emplace_back(worker, S, i);
    return S;
}

int pgeof_server_submit(pgeof_server* S, uint64_t id, const float* xyz, int64_t n, uint32_t* nn_out, float* feat_out)
{
    if (!S || n < 0 || (n > 0 && (!xyz || !feat_out))) return PGEOF_EINVAL;
    if ((int64_t)S->k > n && n > 0) return PGEOF_EINVAL;
    Request r{id, xyz, n, nn_out, feat_out, carve_of(n, S->k).total, 0, Clock::now()};
    {
        std::lock_guard<std::mutex> l(S->mu);
        if (r.footprint > S->budget) return PGEOF_ENOMEM;
        S->queue.push_back(r);
    }
    S->cv_queue.notify_all();
    return PGEOF_OK;
}

int pgeof_server_wait(pgeof_server* S, int max, uint64_t* ids, int* status, double* latency_ms, double timeout_ms)
{
    if (!S || max < 1) return 0;
    std::unique_lock<std::mutex> l(S->mu);
    S->cv_done.wait_for(l, std::chrono::duration<double, std::milli>(timeout_ms), [&] { return !S->done.empty(); });
    int c = 0;
    while (c < max && !S->done.empty())
    {
        const Done d = S->done.front();
        S->done.pop_front();
        ids[c] = d.id;
        if (status) status[c] = d.status;
        if (latency_ms) latency_ms[c] = d.latency_ms;
        ++c;
    }
    return c;
}

size_t pgeof_server_footprint(int64_t n, uint32_t knn) { return carve_of(n, knn).total; }

void pgeof_server_destroy(pgeof_server* S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> l(S->mu);
        S->stop = true;
    }
    S->cv_queue.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& s : S->streams)
    {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(S->pool_base);
    delete S;
}

}  // extern "C"

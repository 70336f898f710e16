// Serving loop: point clouds of uneven size arrive as requests; each runs
// knn_search(k) -> CSR -> compute_features on the device
// (pgeof_knn_features_device).
//
// - `n_slots` requests are in flight at once, each on its own HIP stream
//   with its own device workspace, so their kernels and copies overlap on
//   the GPU (continuous batching: a request is admitted as soon as a slot
//   and enough of the HBM budget are free, not when a batch completes).
// - Admission is FIFO by HBM footprint: the head request waits until the
//   sum of the footprints of the requests in flight plus its own fits the
//   budget; a request larger than the whole budget is rejected at submit.
// - The grid and the CSR stay in the slot's device memory; the features (and
//   the CSR when the caller asks for it) are copied straight into the
//   caller's arrays.
// - Each request's result does not depend on what else is in flight: every
//   kernel is deterministic and a slot's buffers are private to it.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pgeof_hip.h"

namespace
{

using Clock = std::chrono::steady_clock;

struct Request
{
    uint64_t id;
    const float* xyz;
    int64_t n;
    uint32_t* nn_out;  // may be null
    float* feat_out;
    size_t footprint;
    Clock::time_point t_submit;
};

struct Done
{
    uint64_t id;
    int status;
    double latency_ms;
};

size_t footprint_of(int64_t n, uint32_t k)
{
    const size_t pts = (size_t)n;
    return pts * 12 + pts * k * 4 + (pts + 1) * 4 + pts * 11 * 4 + pgeof_knn_workspace_size(n, n, k);
}

struct Slot
{
    hipStream_t stream = nullptr;
    char* dev = nullptr;
    size_t dev_bytes = 0;
};

}  // namespace

struct pgeof_server
{
    int device = 0;
    uint32_t k = 20;
    size_t budget = 0;
    size_t in_use = 0;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_queue;  // queue / budget changes
    std::condition_variable cv_done;
    std::deque<Request> queue;
    std::deque<Done> done;
    std::vector<std::thread> workers;
    std::vector<Slot> slots;
    std::string err;
};

namespace
{

int ensure(Slot& s, size_t dev_bytes)
{
    if (dev_bytes > s.dev_bytes)
    {
        if (s.dev) (void)hipFree(s.dev);
        s.dev = nullptr;
        s.dev_bytes = 0;
        if (hipMalloc(&s.dev, dev_bytes) != hipSuccess) return PGEOF_ENOMEM;
        s.dev_bytes = dev_bytes;
    }
    return PGEOF_OK;
}

inline size_t al(size_t b) { return (b + 255) & ~size_t(255); }

int run_request(pgeof_server* S, Slot& slot, const Request& r)
{
    const uint32_t k = S->k;
    const size_t n = (size_t)r.n;
    if (n == 0) return PGEOF_OK;
    const size_t b_x = al(n * 12), b_nn = al(n * k * 4), b_ptr = al((n + 1) * 4), b_f = al(n * 11 * 4);
    const size_t b_ws = pgeof_knn_workspace_size(r.n, r.n, k);
    int rc = ensure(slot, b_x + b_nn + b_ptr + b_f + b_ws + 4096);
    if (rc) return rc;
    char* d = slot.dev;
    float* d_x = (float*)d;
    uint32_t* d_nn = (uint32_t*)(d + b_x);
    uint32_t* d_ptr = (uint32_t*)(d + b_x + b_nn);
    float* d_f = (float*)(d + b_x + b_nn + b_ptr);
    void* d_ws = d + b_x + b_nn + b_ptr + b_f;
    hipStream_t st = slot.stream;
    // straight from and into the caller's memory: the runtime stages pageable
    // memory itself, pipelined (measured faster than a pinned buffer per slot
    // plus host copies)
    if (hipMemcpyAsync(d_x, r.xyz, n * 12, hipMemcpyHostToDevice, st) != hipSuccess) return PGEOF_ERUNTIME;
    rc = pgeof_knn_features_device(d_x, r.n, k, d_nn, d_ptr, nullptr, d_f, d_ws, b_ws, st);
    if (rc) return rc;
    if (hipMemcpyAsync(r.feat_out, d_f, n * 11 * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return PGEOF_ERUNTIME;
    if (r.nn_out && hipMemcpyAsync(r.nn_out, d_nn, n * k * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return PGEOF_ERUNTIME;
    if (hipStreamSynchronize(st) != hipSuccess) return PGEOF_ERUNTIME;
    return PGEOF_OK;
}

void worker(pgeof_server* S, int idx)
{
    (void)hipSetDevice(S->device);
    Slot& slot = S->slots[idx];
    for (;;)
    {
        Request r;
        {
            std::unique_lock<std::mutex> l(S->mu);
            S->cv_queue.wait(l, [&] {
                return S->stop || (!S->queue.empty() && S->in_use + S->queue.front().footprint <= S->budget);
            });
            if (S->stop && S->queue.empty()) return;
            if (S->stop && !(S->in_use + S->queue.front().footprint <= S->budget)) return;
            r = S->queue.front();
            S->queue.pop_front();
            S->in_use += r.footprint;
        }
        S->cv_queue.notify_all();
        const int rc = run_request(S, slot, r);
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - r.t_submit).count();
        {
            std::lock_guard<std::mutex> l(S->mu);
            S->in_use -= r.footprint;
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
    S->budget = hbm_budget_bytes;
    S->slots.resize((size_t)n_slots);
    for (auto& s : S->slots)
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess)
        {
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
    Request r{id, xyz, n, nn_out, feat_out, footprint_of(n, S->k), Clock::now()};
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

size_t pgeof_server_footprint(int64_t n, uint32_t knn) { return footprint_of(n, knn); }

void pgeof_server_destroy(pgeof_server* S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> l(S->mu);
        S->stop = true;
    }
    S->cv_queue.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& s : S->slots)
    {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.dev) (void)hipFree(s.dev);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    delete S;
}

}  // extern "C"

/*
 * pgeof_hip.h -- C ABI of the MI355X (gfx950) implementation of pgeof's hot path.
 *
 * Library: point_geometric_features_amd/libpgeof_hip.so (HIP kernels + host runtime).
 * All functions return PGEOF_OK (0) or an error code; pgeof_last_error() gives
 * the message of the calling thread's last failure.  No function here takes or
 * returns a torch or numpy type: pointers, sizes and a HIP stream (void*).
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * the pgeof repository, drprojects/point_geometric_features).
 */
#ifndef PGEOF_HIP_H
#define PGEOF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGEOF_OK 0
#define PGEOF_EINVAL 1   /* reference throws std::invalid_argument -> Python ValueError */
#define PGEOF_ERUNTIME 2 /* HIP runtime failure (no device, launch failure, ...) */
#define PGEOF_ENOMEM 3   /* device or workspace memory exhausted */

/* Message of the calling thread's last error ("" if none). */
const char* pgeof_last_error(void);

/* ABI version (major * 100 + minor). */
int pgeof_abi_version(void);

/* Number of HIP devices visible to the process (0 without a GPU). */
int pgeof_device_count(void);

/* ------------------------------------------------------------------------ */
/* Host-memory entry points.  Inputs and outputs are host arrays; the call    */
/* stages them to the current HIP device on a per-thread stream and returns   */
/* when the outputs are written.  Thread-safe; the Python binding calls these */
/* with the GIL released.                                                     */
/* ------------------------------------------------------------------------ */

/* Replaces pgeof::nanoflann_knn_search<float> (include/nn_search.hpp:31-67),
 * bound as pgeof.knn_search (src/pgeof_ext.cpp:118).
 * data (n_data,3) f32 row-major; query (n_query,3) f32 (may alias data).
 * indices/sqr_dist: (n_query, knn) outputs, rows sorted by (sqr_dist, index).
 * EINVAL if knn > n_data (nn_search.hpp:37). */
int pgeof_knn_search(const float* data, int64_t n_data, const float* query, int64_t n_query, uint32_t knn,
                     uint32_t* indices, float* sqr_dist);

/* Replaces pgeof::nanoflann_radius_search<float> (include/nn_search.hpp:85-132),
 * bound as pgeof.radius_search (src/pgeof_ext.cpp:130-132).
 * indices (n_query, max_knn) i32 padded with -1; sqr_dist padded with 0 except
 * that a row with fewer than max_knn hits holds r*r in its last column
 * (nanoflann RKNNResultSet::init).  Hits are d^2 < r*r, nearest first.
 * EINVAL if max_knn > n_data (nn_search.hpp:92-95). */
int pgeof_radius_search(const float* data, int64_t n_data, const float* query, int64_t n_query,
                        float search_radius, uint32_t max_knn, int32_t* indices, float* sqr_dist);

/* Replaces pgeof::compute_geometric_features<float,11> (include/pgeof.hpp:75-117),
 * bound as pgeof.compute_features (src/pgeof_ext.cpp:34-36).
 * nn (n_nn,) u32, nn_ptr (n_nn_ptr,) u32 CSR; features (n_nn_ptr-1, 11) f32.
 * EINVAL if k_min < 1 (pgeof.hpp:81), if n_nn_ptr < 1, or if the CSR indexes
 * outside nn or xyz (undefined behaviour in the reference). */
int pgeof_compute_features(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                           const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min, float* features);

/* Replaces pgeof::compute_geometric_features_multiscale<float,11>
 * (include/pgeof.hpp:159-211), bound as pgeof.compute_features_multiscale
 * (src/pgeof_ext.cpp:59-61).  features (n_nn_ptr-1, n_scales, 11) f32.
 * EINVAL unless k_scales is non-decreasing with every entry >= 1 (pgeof.hpp:123-132, 165-168). */
int pgeof_compute_features_multiscale(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                      const uint32_t* nn_ptr, int64_t n_nn_ptr, const uint32_t* k_scales,
                                      int64_t n_scales, float* features);

/* Replaces pgeof::compute_geometric_features_optimal<float,12>
 * (include/pgeof.hpp:243-310), bound as pgeof.compute_features_optimal
 * (src/pgeof_ext.cpp:86-89).  features (n_nn_ptr-1, 12) f32, column 11 = k_opt.
 * EINVAL if k_min < 1 && k_min_search < 1 (pgeof.hpp:250), or k_step < 1
 * (a modulo by zero in the reference). */
int pgeof_compute_features_optimal(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                   const uint32_t* nn_ptr, int64_t n_nn_ptr, uint32_t k_min, uint32_t k_step,
                                   uint32_t k_min_search, float* features);

/* Replace pgeof::compute_geometric_features_selected<float|double>
 * (include/pgeof.hpp:325-375, include/pca.hpp:212-295), bound as the two
 * overloads of pgeof.compute_features_selected (src/pgeof_ext.cpp:148-177).
 * selected: EFeatureID values (include/pca.hpp:47-63); features (n, n_selected). */
int pgeof_compute_features_selected_f32(const float* xyz, int64_t n, float search_radius, uint32_t max_knn,
                                        const int32_t* selected, int64_t n_selected, float* features);
int pgeof_compute_features_selected_f64(const double* xyz, int64_t n, double search_radius, uint32_t max_knn,
                                        const int32_t* selected, int64_t n_selected, double* features);

/* ------------------------------------------------------------------------ */
/* Device-memory entry points.  All pointers are device pointers on the       */
/* current HIP device; work is ordered on `stream` (a hipStream_t, NULL = the */
/* null stream).  `workspace` may be NULL: the call then borrows a buffer of  */
/* the library's scratch cache (allocated once, reused once the stream that   */
/* last used it has passed an event recorded after its launches).  The searches */
/* size their grids on the device and never synchronise `stream`: given a     */
/* workspace (and, for pgeof_radius_csr_device, nnz_out = NULL) a call only   */
/* queues kernels, so it can be captured into a HIP graph.                    */
/* ------------------------------------------------------------------------ */

/* Workspace bytes for the search entry points below. */
size_t pgeof_knn_workspace_size(int64_t n_data, int64_t n_query, uint32_t knn);
size_t pgeof_radius_workspace_size(int64_t n_data, int64_t n_query, uint32_t max_knn);

/* Device form of pgeof_knn_search.  sqr_dist may be NULL. */
int pgeof_knn_search_device(const float* data, int64_t n_data, const float* query, int64_t n_query, uint32_t knn,
                            uint32_t* indices, float* sqr_dist, void* workspace, size_t workspace_bytes,
                            void* stream);

/* Device form of pgeof_radius_search. */
int pgeof_radius_search_device(const float* data, int64_t n_data, const float* query, int64_t n_query,
                               float search_radius, uint32_t max_knn, int32_t* indices, float* sqr_dist,
                               void* workspace, size_t workspace_bytes, void* stream);

/* knn_search(xyz, xyz, knn) -> CSR (README.md:136-141: nn = idx.flatten(),
 * nn_ptr = arange(n+1)*knn) -> compute_features(xyz, nn, nn_ptr, k_min=1),
 * fused: the features are computed from the search's final neighbour lists
 * without re-reading nn.  nn (n*knn) u32, nn_ptr (n+1) u32, features (n,11);
 * sqr_dist (n*knn) may be NULL.  Results equal the two-step path bit for bit. */
int pgeof_knn_features_device(const float* xyz, int64_t n, uint32_t knn, uint32_t* nn, uint32_t* nn_ptr,
                              float* sqr_dist, float* features, void* workspace, size_t workspace_bytes,
                              void* stream);

/* radius_search(xyz, xyz, r, max_knn) -> CSR of the hits (README.md:158-159:
 * nn_ptr = [0, cumsum(count)], nn = idx[idx >= 0]).  nn must hold n*max_knn
 * entries; nn_ptr (n+1).  If nnz_out (host) is non-NULL the call synchronises
 * `stream` and stores nn_ptr[n] there; with NULL it stays asynchronous. */
size_t pgeof_radius_csr_workspace_size(int64_t n, uint32_t max_knn);
int pgeof_radius_csr_device(const float* xyz, int64_t n, float search_radius, uint32_t max_knn, uint32_t* nn,
                            uint32_t* nn_ptr, int64_t* nnz_out, void* workspace, size_t workspace_bytes,
                            void* stream);

/* radius_search(xyz, xyz, r, max_knn) -> CSR (as pgeof_radius_csr_device) ->
 * compute_features_multiscale(xyz, nn, nn_ptr, k_scales)
 * (include/pgeof.hpp:159-211, README.md:158-159) in one call, with no
 * read-back: capturable into a HIP graph when a workspace of
 * pgeof_radius_multiscale_workspace_size bytes is given.  The features come
 * from a CSR kernel after the search, or (tuning knob "fuse_multiscale") from
 * the search kernel itself, off each row's sorted hits.  k_scales is a host
 * array (its values travel as kernel arguments).  nn (n*max_knn) u32,
 * nn_ptr (n+1) u32, features (n, n_scales, 11) f32, equal bit for bit to the
 * two-call path.  EINVAL as pgeof_radius_csr_device and
 * pgeof_compute_features_multiscale. */
size_t pgeof_radius_multiscale_workspace_size(int64_t n, uint32_t max_knn, int64_t n_scales);
int pgeof_radius_multiscale_device(const float* xyz, int64_t n, float search_radius, uint32_t max_knn,
                                   const uint32_t* k_scales, int64_t n_scales, uint32_t* nn, uint32_t* nn_ptr,
                                   float* features, void* workspace, size_t workspace_bytes, void* stream);

/* knn_search(xyz, xyz, knn) -> CSR (nn_ptr = arange*knn) ->
 * compute_features_optimal(xyz, nn, nn_ptr, k_min, k_step, k_min_search)
 * (include/pgeof.hpp:243-310, README.md:136-141) in one call: the feature
 * kernel visits the rows in the search's spatial order.  Workspace:
 * pgeof_knn_workspace_size(n, n, knn) + 256 bytes; with it the call only
 * queues kernels (capturable into a HIP graph).  features (n, 12) f32, equal
 * bit for bit to the two-call path.  EINVAL as pgeof_knn_features_device and
 * pgeof_compute_features_optimal. */
int pgeof_knn_optimal_device(const float* xyz, int64_t n, uint32_t knn, uint32_t k_min, uint32_t k_step,
                             uint32_t k_min_search, uint32_t* nn, uint32_t* nn_ptr, float* features, void* workspace,
                             size_t workspace_bytes, void* stream);

/* Device forms of the CSR feature drivers.  If err is non-NULL (a device u32)
 * the call is fully asynchronous and ORs 1 (bad neighbour index) / 2 (bad
 * nn_ptr) into *err; if NULL the call synchronises `stream` and returns
 * PGEOF_EINVAL on such input.  Argument checks are those of the host forms. */
int pgeof_compute_features_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                  const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min, float* features,
                                  uint32_t* err, void* stream);
int pgeof_compute_features_multiscale_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                             const uint32_t* nn_ptr, int64_t n_nn_ptr,
                                             const uint32_t* k_scales_host, int64_t n_scales, float* features,
                                             uint32_t* err, void* stream);
int pgeof_compute_features_optimal_device(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                          const uint32_t* nn_ptr, int64_t n_nn_ptr, uint32_t k_min,
                                          uint32_t k_step, uint32_t k_min_search, float* features, uint32_t* err,
                                          void* stream);

/* Largest CSR (rows = n_nn_ptr - 1) that pgeof_compute_features_device_noalloc
 * accepts. */
#define PGEOF_NOALLOC_MAX_ROWS 4194304

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pgeof_compute_features_device with its scratch (the
 * spatial row order of inputs of 65536 rows or more) taken from a 60 MiB
 * array that the library holds on each device it runs on until the process
 * ends.  It calls no allocator and never synchronises: err (a device u32,
 * zeroed by the caller) is required and receives the error bits.  At most
 * PGEOF_NOALLOC_MAX_ROWS rows (EINVAL beyond).  Calls on one device share
 * the array: calls that overlap in time (different streams, no ordering
 * between them) race; order them on one stream.  Results equal those of
 * pgeof_compute_features_device bit for bit. */
int pgeof_compute_features_device_noalloc(const float* xyz, int64_t n_xyz, const uint32_t* nn, int64_t n_nn,
                                          const uint32_t* nn_ptr, int64_t n_nn_ptr, uint64_t k_min,
                                          float* features, uint32_t* err, void* stream);

/* ------------------------------------------------------------------------ */
/* Serving loop (no counterpart in the reference; the reference processes one  */
/* cloud per call on the CPU).  Requests are point clouds of any size; each    */
/* runs pgeof_knn_features_device with k = knn on its own stream; up to        */
/* n_slots run concurrently.  The server's device memory is one pool of        */
/* hbm_budget_bytes allocated at creation; a request is admitted (FIFO) when   */
/* a contiguous range of its footprint (pgeof_server_footprint) is free, and   */
/* returns it on completion: no allocation happens while serving.             */
/* ------------------------------------------------------------------------ */
typedef struct pgeof_server pgeof_server;

/* Create on the current device (allocates the pool); NULL on failure. */
pgeof_server* pgeof_server_create(int n_slots, size_t hbm_budget_bytes, uint32_t knn);

/* Enqueue a request.  xyz (n,3) f32 host; feat_out (n,11) f32 host;
 * nn_out (n*knn) u32 host or NULL.  The buffers must stay valid until the
 * request is reported by pgeof_server_wait.  PGEOF_ENOMEM if the request
 * alone exceeds the budget, PGEOF_EINVAL if knn > n. */
int pgeof_server_submit(pgeof_server* server, uint64_t id, const float* xyz, int64_t n, uint32_t* nn_out,
                        float* feat_out);

/* Wait up to timeout_ms for completed requests; writes up to max entries of
 * (id, status, latency from submit to completion in ms); returns the count. */
int pgeof_server_wait(pgeof_server* server, int max, uint64_t* ids, int* status, double* latency_ms,
                      double timeout_ms);

/* HBM bytes a request of n points occupies while in flight. */
size_t pgeof_server_footprint(int64_t n, uint32_t knn);

/* Stop the workers (requests still queued are dropped) and free the server. */
void pgeof_server_destroy(pgeof_server* server);

/* ------------------------------------------------------------------------ */
/* Tuning and profiling (no counterpart in the reference).                   */
/* ------------------------------------------------------------------------ */

/* Set a tuning knob: "knn_occupancy" (target points per grid cell, as a
 * fraction of k), "knn_cell_points_max" (cap of that target), "grid_cells_per_point" (cap on grid cells / points),
 * "lds_budget" (bytes of LDS per workgroup for the search heaps),
 * "radius_cell_factor" (grid cell edge / search radius), "grid_levels"
 * (depth of the kNN grid pyramid, 1..4), "near_first" (1: cell rows are
 * visited nearest-first, 0: in index order), "budget_cells" (budgeted kNN
 * pass: a query whose ball reaches farther than this many cells from its own
 * cell is deferred to the next pass; 1..64), "budget_shells" (budgeted kNN
 * pass: cubic shells around the query's cell scanned before a query that has
 * not found k candidates is deferred; 1..budget_cells), "budget_levels"
 * (pyramid levels, finest first, that get a budgeted pass; coarser levels
 * serve the unbudgeted pass only), "grid_level_ratio" (cell edge
 * ratio between consecutive pyramid levels, 2..8), "csr_order" (1: the
 * plain and optimal-k kernels visit the rows of inputs of 65536 rows or more
 * in the spatial order of their points, 0: in index order), "heap_tail"
 * (1: search heaps of k > 21 keep nodes 21..k-1 in global memory and only
 * nodes 5..20 in LDS; searches of over 2 M queries keep those tails per
 * hardware wave slot, 128 MiB per 4 tail nodes, smaller ones per query;
 * 2: per hardware wave slot at any size; 0: the whole heap in LDS while it
 * fits), "staging"
 * (wave-shared LDS staging of cell rows: 0 never (default: each lane streams
 * its rows' candidates from L2), 1 radius search and, for k >= 32, the
 * budgeted kNN pass, 2 radius search and the budgeted kNN pass at any k), "fuse_multiscale" (pgeof_radius_multiscale_device: 1 the
 * features in the search kernel, 0 in a CSR kernel after it).
 * Knobs change speed only, never results. Values must be finite and >= 0. */
int pgeof_set_tuning(const char* key, double value);
double pgeof_get_tuning(const char* key);

/* Autotuner.  Times the workload on the given cloud (device pointer xyz, n
 * points) with the current knobs and with a sweep of the knobs that move its
 * speed (grid resolution, block size through the heaps' LDS budget, row
 * staging, and for "radius_multiscale" whether the search kernel computes the
 * features), `reps` timed runs each (median), for kNN the fastest of those
 * also with the other pyramid depth (one level or three), then the four
 * fastest again with 2 * reps + 1 runs each, taken in turns, on `stream`.
 * The set with the lowest median is kept for the workload's search family ("knn_features" and "knn_optimal":
 * kNN, knn = k; "radius_multiscale": radius, max_knn = k, scales k_scales)
 * and its size class (ceil(log2 n)): later calls with that k and size class
 * use it, workspace sizes included, until pgeof_autotune_clear.  It allocates
 * its own buffers and synchronises: run it before serving, not inside a
 * graph.  report (if non-NULL) receives a JSON summary (times in ms of every
 * candidate, the chosen knob set).  Knobs change speed only, never results. */
int pgeof_autotune(const char* workload, const float* xyz, int64_t n, uint32_t k, float radius,
                   const uint32_t* k_scales, int64_t n_scales, int reps, void* stream, char* report,
                   size_t report_len);
void pgeof_autotune_clear(void);

/* When enabled, every kernel launch is bracketed by HIP events on its stream.
 * pgeof_profile_report writes {"kernel": [launches, total_ms], ...} as JSON
 * into buf (synchronising the recorded events) and returns the length. */
void pgeof_profile_enable(int on);
int pgeof_profile_report(char* buf, size_t len);
void pgeof_profile_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* PGEOF_HIP_H */

/*
 * adaptive_trees.h -- C ABI of libadaptive_trees_hip.so, the MI355X (gfx950)
 * adaptive octree / binary-tree build and update engine.
 *
 * Every entry point is a plain C function over plain pointers and sizes, so
 * any FFI (ctypes, cgo, JNI, N-API) can bind it. Coordinates are DEVICE
 * pointers owned by the caller, as in the reference's constructors
 * (include/adaptive_octree/octree.hpp:36-43, kdtree3d.hpp:28-35): they are not
 * copied, must outlive the tree and are re-read by every update().
 *
 * Status codes replace the reference's exceptions one to one
 * (include/adaptive_octree/tree.hpp:110-130, src/octree/octree.cu:326-330,
 * 387-410, 474-475, 187-189): AT_EINVAL <-> std::invalid_argument,
 * AT_ELOGIC <-> std::logic_error, AT_ELENGTH <-> std::length_error. A HIP
 * failure returns AT_EHIP instead of the reference's exit(EXIT_FAILURE)
 * (include/adaptive_octree/config.hpp:164-172). at_last_error() returns the
 * message of the last failing call on the calling thread.
 *
 * Threading: one tree is driven by one host thread at a time, on the stream
 * named in its config (NULL = the default stream). at_tree_update() and
 * at_tree_compute_views() launch only asynchronous work and may be recorded
 * into a HIP graph; build, needs_resize and resize_buffers synchronise.
 */
#ifndef ADAPTIVE_TREES_H
#define ADAPTIVE_TREES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    AT_OK = 0,
    AT_EINVAL = 1,  /* std::invalid_argument in the reference */
    AT_ELOGIC = 2,  /* std::logic_error */
    AT_ELENGTH = 3, /* std::length_error */
    AT_EHIP = 4,    /* HIP runtime failure */
    AT_ENOMEM = 5,  /* device or host allocation failure */
    AT_EIO = 6      /* std::runtime_error of the particle-file readers */
};

/* ---- enums (include/adaptive_octree/config.hpp:48-50) ------------------ */
enum { AT_TREE_OCTREE = 0, AT_TREE_BINARY = 1 };
enum { AT_LEAF_COUNT = 0, AT_NF_COUNT = 1 };
enum { AT_MORTON = 0, AT_HILBERT = 1 };

/* Box<Real> (config.hpp:95-101). Values are converted to the tree's Real. */
typedef struct at_box {
    double xmin, xmax, ymin, ymax, zmin, zmax;
} at_box;

/*
 * Tree configuration: TreeConfig<Real> (config.hpp:143-155) with the backend
 * defaults of Tree::backendConfig (tree.hpp:108-133).
 *   max_depth          -1 = backend default (octree: 21 build passes, which
 *                      the reference does not enforce as a depth,
 *                      config.hpp:125; binary: depth 63).
 *   update_iterations  -1 = backend default (octree 2, binary 6).
 *   views              {treeStructure, particleCountsPerBox, particleMapping}
 *   stream             hipStream_t of the caller, NULL = default stream.
 */
typedef struct at_config {
    int32_t max_depth;
    uint32_t max_particles_per_leaf;
    int32_t update_iterations;
    int32_t split_criterion;
    int32_t sfc_kind;
    uint8_t views[3];
    uint8_t reserved0;
    at_box box;
    void *stream;
} at_config;

/* Output buffers; names follow the getters of tree.hpp:62-103. */
enum {
    AT_BUF_PARTICLE_KEYS = 0,       /* P,    uint64[numParticles]          */
    AT_BUF_PERM = 1,                /* Perm, uint32[numParticles]          */
    AT_BUF_CORNERSTONE = 2,         /* K,    uint64[numLeaves + 1]         */
    AT_BUF_LEAF_COUNTS = 3,         /* N,    uint32[numLeaves]             */
    AT_BUF_NEAR_FIELD_COUNTS = 4,   /* NF,   uint32[numLeaves] (octree)    */
    AT_BUF_SFC_BOX_INDEX = 5,       /* uint64[numNodes]                    */
    AT_BUF_HAS_BOX_SPLIT = 6,       /* uint8[numNodes]                     */
    AT_BUF_BOX_DEPTH = 7,           /* uint8[numNodes]                     */
    AT_BUF_PARENT_INDEX = 8,        /* int32[numNodes]                     */
    AT_BUF_CHILD_INDEX = 9,         /* int32[numNodes]                     */
    AT_BUF_LEVEL_OFFSET = 10,       /* int32[23] octree, int32[65] binary  */
    AT_BUF_PARTICLE_BEGIN_INDEX = 11, /* uint32[numNodes]                  */
    AT_BUF_PARTICLE_COUNTS = 12,    /* uint32[numNodes + 1]                */
    AT_BUF_SFC_PARTICLE_INDEX = 13, /* uint64[numActiveParticles]          */
    AT_BUF_COUNT = 14
};

/* Host-visible sizes from the last build()/needs_resize() (octree.hpp:60-69). */
typedef struct at_scalars {
    uint64_t num_particles;
    uint64_t num_active_particles;
    uint64_t num_nodes;
    uint64_t num_leaves;
    uint64_t leaf_capacity;
    int32_t tree_type;
    int32_t split_criterion;
    int32_t sfc_kind;
    int32_t real_bytes;
    int32_t active_buffer;
    int32_t update_iterations;
    int32_t max_depth;
    uint32_t max_particles_per_leaf;
} at_scalars;

typedef struct at_tree at_tree;

/* Constructor: Octree(x,y,z,n,config) (octree.cu:322-425), KDTree3D(...)
 * (kdtree3d.cu:184-219), Tree(...) (tree.hpp:24-27). real_bytes is 4 or 8. */
int at_tree_create(int real_bytes, int tree_type, const void *x, const void *y,
                   const void *z, size_t count, const at_config *config,
                   at_tree **out);
int at_tree_destroy(at_tree *tree);
/* Octree<Real>::build (octree.cu:428-469), KDTree3D<Real>::build (kdtree3d.cu:227-257). */
int at_tree_build(at_tree *tree);
/* update() (octree.cu:472-481, kdtree3d.cu:259-270): asynchronous, capturable. */
int at_tree_update(at_tree *tree);
/* computeViews(ViewConfig) (octree.cu:484-497); views == NULL uses config.views. */
int at_tree_compute_views(at_tree *tree, const uint8_t *views);
/* doBuffersNeedResize() (octree.cu:499-537, kdtree3d.cu:283-287). */
int at_tree_needs_resize(at_tree *tree, int *needs_resize);
/* resizeBuffers() (octree.cu:540-551, kdtree3d.cu:289-297). */
int at_tree_resize_buffers(at_tree *tree);
/* Getter of one output buffer: device address, logical size, capacity (elements). */
int at_tree_buffer(at_tree *tree, int buffer_id, const void **data, size_t *size,
                   size_t *capacity);
int at_tree_scalars(at_tree *tree, at_scalars *out);
/* numLeaves_d(), numNodes_d(), numActiveParticles_d() (octree.hpp:62-66). */
int at_tree_device_counters(at_tree *tree, const int32_t **num_leaves,
                            const int32_t **num_nodes, const uint32_t **num_active);
/* Measurement (no reference counterpart): the passes of the last update()
 * that changed the tree, i.e. ran before it converged (a pass rejected for
 * capacity does not count). Synchronises the tree's stream. */
int at_tree_update_passes(at_tree *tree, uint32_t *passes);
/* maxAchievedDepth() (octree.cu:666-670, kdtree3d.cu:404-415); synchronises. */
int at_tree_max_achieved_depth(at_tree *tree, uint32_t *depth);
/* Leaf-count convergence check of the reference's harness
 * (benchmarks/drivers/cstone_check.cu:17-43) on the tree as its getters see
 * it (call doBuffersNeedResize after update), populations recounted from P:
 * *overfull = leaves with more than `bucket` particles and a key range
 * above 1; *mergeable = aligned groups of sibling leaves (8 for an octree;
 * for a binary tree the analogous 2) that exactly tile their parent and
 * hold at most `bucket` particles together. A converged LeafCount tree has
 * neither (unless leaves reach the finest level); a fixed-pass update may
 * leave some. Synchronises. */
int at_tree_leaf_count_violations(at_tree *tree, uint32_t bucket, uint64_t *overfull, uint64_t *mergeable);
/* print(std::ostream&) (src/octree/octree_print.cu:57-382,
 * src/kdtree3d/kdtree3d_print.cu:11-43): the same text, formatted as on a
 * stream in its default state. Copies at most len-1 characters and a NUL
 * into buf (may be NULL with len 0); *needed (may be NULL) receives the
 * full length + 1. Synchronises. */
int at_tree_print(at_tree *tree, char *buf, size_t len, size_t *needed);
/* Last error message of the calling thread ("" if none). */
const char *at_last_error(void);

/* ---- device plumbing for FFI callers without their own HIP bindings ---- */
int at_device_count(int *count);
int at_device_malloc(void **ptr, size_t bytes);
int at_device_free(void *ptr);
/* kind: 1 host->device, 2 device->host, 3 device->device; synchronous. */
int at_memcpy(void *dst, const void *src, size_t bytes, int kind);
int at_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
int at_stream_create(void **stream);
int at_stream_destroy(void *stream);
int at_stream_synchronize(void *stream);
int at_device_synchronize(void);
int at_event_create(void **event);
int at_event_destroy(void *event);
int at_event_record(void *event, void *stream);
int at_event_elapsed_ms(void *start, void *stop, float *ms);
/* Record tree->update() on `stream` into an executable HIP graph (the
 * reference's capture protocol, tests/tree_test.cu:44-63). */
int at_graph_capture_update(at_tree *tree, void *stream, void **graph_exec);
int at_graph_launch(void *graph_exec, void *stream);
int at_graph_destroy(void *graph_exec);

/* ---- measurement ------------------------------------------------------ */
/* Event timing of labelled launches ("keygen", "sort_pass", "decide_scan",
 * "leaf_counts", "near_field", "bin_direct", "views", "copy") outside graph
 * capture; query returns the summed milliseconds and the launch count. */
int at_profile_enable(int on);
int at_profile_reset(void);
int at_profile_query(const char *label, double *total_ms, unsigned long long *launches);
/* Streaming device copy of `bytes` (multiple of 16): the known-traffic
 * kernel that calibrates hardware byte counters. */
int at_bench_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ---- batches: many independent octrees, shared launches ----------------
 * Batched build/update of independent trees (no reference counterpart: the
 * reference drives one Tree per call; this replaces a loop of
 * Tree<Real,Octree>::build()/update() calls, tree.hpp:35-41, with one
 * segmented launch per stage). All members share `config` (octree only;
 * LeafCount or NFCount, Morton or Hilbert); member k borrows the device
 * coordinates x[k], y[k], z[k] of counts[k] particles. Each member's
 * results equal those of its own Tree after the same calls.
 * at_batch_update() is capturable; at_batch_graph_capture_update() records
 * it on the batch's (non-NULL) stream. Member buffers are read with
 * at_batch_tree_buffer (same ids as at_tree_buffer). */
typedef struct at_batch at_batch;
int at_batch_create(int real_bytes, int tree_type, size_t num_trees, const void *const *x,
                    const void *const *y, const void *const *z, const size_t *counts,
                    const at_config *config, at_batch **out);
int at_batch_destroy(at_batch *batch);
int at_batch_build(at_batch *batch);
int at_batch_update(at_batch *batch);
/* flags: num_trees ints (may be NULL); count: members needing a resize. */
int at_batch_needs_resize(at_batch *batch, int *flags, size_t *count);
/* Grows the members whose last update was rejected for capacity. */
int at_batch_resize_buffers(at_batch *batch);
int at_batch_size(at_batch *batch, size_t *num_trees);
int at_batch_tree_buffer(at_batch *batch, size_t tree, int buffer_id, const void **data,
                         size_t *size, size_t *capacity);
int at_batch_tree_scalars(at_batch *batch, size_t tree, at_scalars *out);
int at_batch_graph_capture_update(at_batch *batch, void *stream, void **graph_exec);

/* ---- serving: resident trees, admission by HBM budget, continuous batching
 * No reference counterpart (the reference has no request loop); for callers
 * that issue many independent build/update requests and want them to
 * overlap on one GPU. Requests return at once with an id; at_server_wait
 * blocks until one is done. A tree's requests run in submission order; an
 * update may bring new positions (device arrays copied into the tree's
 * coordinates before its update) or NULLs when they were changed in place
 * (only safe when no request of that tree is in flight). Build requests are
 * admitted in arrival order while their estimated bytes fit hbm_budget
 * (a release frees them). Each worker runs its queued requests in rounds:
 * up to max_batch requests, at most one per tree and in each tree's order;
 * the round's octree builds, then its octree updates, of one configuration
 * and up to 4M particles per tree share their launches (16M particles per
 * group); other requests run tree by tree; the round synchronises once
 * after its updates. Tree buffers may be read once the tree's last request
 * is done. */
typedef struct at_server at_server;
typedef struct at_server_config {
    uint64_t hbm_budget_bytes;
    int32_t workers;       /* worker threads, one HIP stream each (default 4) */
    int32_t max_batch;     /* requests per worker round (default 16) */
    int32_t update_graphs; /* nonzero: replay each tree's update from a HIP graph,
                              captured at its first update (trees then run
                              one by one, no shared launches); while a
                              capture runs, HIP fails legacy-default-stream
                              operations in every thread of the process */
    int32_t reserved0;
} at_server_config;
typedef struct at_server_stats {
    uint64_t bytes_in_use, bytes_peak, hbm_budget;
    uint64_t resident_trees, waiting_builds, completed;
    uint64_t batched_updates, update_batches; /* updates in rounds of several updates; such rounds */
    uint64_t grouped_builds, grouped_updates; /* requests that shared launches with other trees */
} at_server_stats;
int at_server_create(const at_server_config *config, at_server **out);
/* Finishes the requests already given to a worker first; builds still waiting
 * for budget, and requests held behind them, complete with AT_ELOGIC (threads
 * blocked in at_server_wait on them return). */
int at_server_destroy(at_server *server);
int at_server_submit_build(at_server *server, int real_bytes, int tree_type, const void *x,
                           const void *y, const void *z, size_t count, const at_config *config,
                           uint64_t *tree_id, uint64_t *request_id);
int at_server_submit_update(at_server *server, uint64_t tree_id, const void *new_x,
                            const void *new_y, const void *new_z, uint64_t *request_id);
int at_server_submit_release(at_server *server, uint64_t tree_id, uint64_t *request_id);
/* *status receives the request's own status code (AT_OK on success). */
int at_server_wait(at_server *server, uint64_t request_id, int *status);
int at_server_poll(at_server *server, uint64_t request_id, int *done);
/* Milliseconds since server creation; -1 while not started / not done. */
int at_server_request_times(at_server *server, uint64_t request_id, double *submit_ms,
                            double *start_ms, double *done_ms);
int at_server_tree_buffer(at_server *server, uint64_t tree_id, int buffer_id, const void **data,
                          size_t *size, size_t *capacity);
int at_server_tree_scalars(at_server *server, uint64_t tree_id, at_scalars *out);
int at_server_get_stats(at_server *server, at_server_stats *out);

/* ---- tuning (process-wide; set before building trees) ------------------ */
/* "sort_threads": threads per radix-sort block (256, 512 or 1024);
 * "sort_items":  keys per thread of a radix-sort tile (8, 12 or 16);
 * "sort_ranges": target number of key ranges per sort pass (64..65536);
 * "sort_digit_array": 1 = each scatter pass writes the next digit byte per
 *                     key for the next histogram, 0 = the histogram reads
 *                     the keys;
 * "sort_prefetch":    1 = the scatter loads the next tile while it works on
 *                     the current one (tiles 256/512 x 8/12);
 * "sort_msd":         1 = sorts of at least sort_msd_min_keys (key, value)
 *                     pairs partition by the high digits first and finish
 *                     buckets of up to 2048 keys in LDS, 0 = LSD passes.
 * "sort_msd_min_keys": that threshold (2048 .. 2^31 - 1; default 4096).
 * "sort_digit_bits":  digit width of the LSD passes of sorts below 2^20 keys
 *                     (with sort_msd 0, or below the MSD threshold): 8
 *                     (eight passes) or 11 (six passes of 2048 digits).
 * A tile (sort_threads x sort_items) holds at most 8192 keys; a larger one
 * is rejected when a sort is planned. */
int at_tune_set(const char *key, int value);
int at_tune_get(const char *key, int *value);
/* Autotuner: times the tile shapes, then the range counts, then the
 * histogram source on a sort of n random keys (on `stream`), then the LSD
 * digit width on a sort of min(n, 2^20 - 1) keys; keeps the fastest
 * settings and writes them as JSON into report (may be NULL). Every setting
 * gives identical results; only the schedule changes. Replaces the fixed
 * digit width and tile shape of CUB's SortPairs at src/octree/octree.cu:301-304. */
int at_tune_auto(size_t n, void *stream, char *report, size_t report_len);

/* ---- device radix sort ---------------------------------------------------
 * Stable sort of n 64-bit keys with 32-bit values, in place, on `stream`
 * (device pointers; returns after the stream has finished). slot_values != 0:
 * vals is output only and receives each key's input position 0..n-1 (the
 * permutation the trees store in Perm). The engine behind every tree's
 * key sort (no reference counterpart: the reference calls
 * cub::DeviceRadixSort::SortPairs). n < 2^32. */
int at_sort_pairs(void *keys, void *vals, size_t n, int slot_values, void *stream);
/* The same sort with nothing allocated: no counterpart in the reference (nor
 * an interface it replaces); for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Its workspace is a fixed device array of the library, 14.1 MiB
 * on each GPU it runs on, held until the process ends, so n is limited to
 * 2^20 pairs (larger n fails with AT_ELENGTH before any launch). Results
 * equal at_sort_pairs. Asynchronous: returns once the work is queued on
 * `stream`; synchronise before reading keys/vals. Every call on one GPU
 * shares the array, so calls that overlap on one GPU (other streams or
 * threads) race; calls on one stream are ordered. */
int at_sort_pairs_noalloc(void *keys, void *vals, size_t n, int slot_values, void *stream);

/* ---- particle input formats (host memory) ----------------------------------
 * Text particle files of the reference's tests (tests/particle_io.hpp:35-139,
 * particle_io::readColumns): `columns` (1..16) numbers per row as strtod
 * reads them, each followed by whitespace, '#' or the end of the row; blank
 * rows and rows starting with '#' are skipped; a row "ESC" keeps its slot
 * with every column set to DBL_MAX (the escaped-particle marker,
 * include/adaptive_octree/config.hpp:103-113). On success *data is a
 * library-allocated array [columns][*rows] (column-major; release with
 * at_io_free). A malformed row fails with AT_EIO and names path:line. The
 * x, y, z of particle_io::readParticles are columns 1..3 ("charge x y z")
 * or 0..2 ("x y z charge") of a four-column read. */
int at_io_read_columns(const char *path, int columns, double **data, size_t *rows);
/* One frame of the benchmark harness (benchmarks/drivers/common.hpp:124-170,
 * bench::readFrame): format "cxyz" / "xyzc" = text rows of four numbers
 * (from_chars, one leading '+' allowed) or rows starting with "ESC"; format
 * "bin" = x y z doubles per slot, NaN x = escaped. *xyz = [3][*n] doubles
 * (release with at_io_free). */
int at_io_read_frame(const char *path, const char *format, double **xyz, size_t *n);
void at_io_free(void *p);
/* Harness manifest (benchmarks/drivers/common.hpp:62-104, bench::readManifest):
 * "dataset NAME", "slots N", "box xmin xmax ymin ymax zmin zmax" and
 * "frame INDEX ID FORMAT PATH" lines (PATH relative to the manifest's
 * directory). Fails (AT_EIO) without frames or with slots == 0. Strings are
 * copied NUL-terminated and truncated to the buffer lengths given. */
int at_io_manifest_open(const char *path, void **manifest);
int at_io_manifest_info(const void *manifest, char *dataset, size_t dataset_len, size_t *slots, double *box,
                        size_t *frames);
int at_io_manifest_frame(const void *manifest, size_t i, int *index, long *id, char *format, size_t format_len,
                         char *path, size_t path_len);
void at_io_manifest_close(void *manifest);

/* ---- synthetic workloads (host memory) -------------------------------------
 * No counterpart in the reference, which reads recorded datasets: the seeded
 * inputs of the benchmark workloads. One xoshiro256** stream per call, seeded
 * by splitmix64(seed); uniform doubles (u >> 11) * 2^-53, Gaussians by
 * Box-Muller in double. Coordinates column-major in xyz[3 n]; box =
 * {xmin, xmax, ymin, ymax, zmin, zmax}. The draw order is documented in
 * adaptive_trees_amd/csrc/workloads.cpp. */
int at_workload_uniform(uint64_t seed, size_t n, const double *box, double *xyz);
/* Plummer sphere of scale radius a around (0.5, 0.5, 0.5), truncated at 20 a,
 * clipped to the unit box. */
int at_workload_plummer(uint64_t seed, size_t n, double a, double *xyz);
/* 40% / 40% isotropic Gaussians (sigma = sigma_of_long_side * (xmax - xmin))
 * at 1/4 and 3/4 of x, 20% uniform, clipped to the box. */
int at_workload_two_clusters(uint64_t seed, size_t n, const double *box, double sigma_of_long_side, double *xyz);
/* out = in + N(0, (sigma_of_side * side)^2) per axis, reflected at the box
 * walls and clamped to it. */
int at_workload_reflect_step(uint64_t seed, size_t n, const double *box, double sigma_of_side, const double *in,
                             double *out);
int at_workload_log_uniform_sizes(uint64_t seed, size_t count, double lo, double hi, int64_t *out);
int at_workload_bits(uint64_t seed, size_t count, uint64_t *out); /* the raw stream */

#ifdef __cplusplus
}
#endif

#endif /* ADAPTIVE_TREES_H */

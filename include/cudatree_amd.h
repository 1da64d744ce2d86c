/*
 * cudatree_amd.h -- C-ABI of the MI355X random-forest library
 * (libcudatree_amd.so, HIP for gfx950).
 *
 * The reference (EasonLiao/CudaTree) has no C interface: its boundary is the
 * Python class RandomForestClassifier (cudatree/random_forest.py:45-51, 292,
 * 339, 364, 379) driving PyCUDA kernel launches (prepared_call on templated
 * .cu modules, cudatree/util.py:24-37, random_forest.py:382-460).  This header
 * is the boundary the Python facade (cudatree_amd/random_forest.py, loaded
 * with ctypes) binds in place of those launches; each entry point names the
 * reference code it replaces.
 *
 * Conventions
 *   - Every call returns an int status (CTA_OK == 0); cta_last_error() gives
 *     a thread-local message for the last failure.  No C++ exception crosses
 *     the ABI.
 *   - Host arrays are copied in and out; the library owns its device memory.
 *     The *_device variants take device pointers (e.g. PyTorch-ROCm tensors)
 *     that must stay valid until the call returns.
 *   - A context owns one HIP stream on one device.  Calls on one context
 *     are serialised by its internal mutex; use one context per thread for
 *     concurrency.
 *   - Labels are compact: y[i] in [0, K).  Up to 4,194,303 samples (exact
 *     64-bit Gini arithmetic), 65,535 features and 1,024 classes.
 */
#ifndef CUDATREE_AMD_H
#define CUDATREE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTA_OK 0
#define CTA_EINVAL 1      /* bad argument */
#define CTA_ENOMEM 2      /* device or host allocation failed */
#define CTA_EHIP 3        /* HIP runtime error */
#define CTA_EUNSUPPORTED 4 /* size beyond the library's limits */
#define CTA_EINTERNAL 5   /* device-side consistency check failed */

#define CTA_MAX_SAMPLES 4194303
#define CTA_MAX_FEATURES 65535
#define CTA_MAX_CLASSES 1024

/* Fit scheduling modes.  All three build identical trees (nodes are
 * canonically numbered in depth-first pre-order on export). */
#define CTA_MODE_BATCHED 0  /* every level of every tree in one set of launches */
#define CTA_MODE_HYBRID 1   /* per tree: nodes above bfs_threshold one at a time, the rest level-batched
                               (random_tree.py:616-622, 309-458) */
#define CTA_MODE_PER_NODE 2 /* every node in its own launch sequence, trees in series
                               (random_tree.py:580-676 with bfs_threshold = 0) */

typedef struct cta_ctx cta_ctx;
typedef struct cta_dataset cta_dataset;
typedef struct cta_forest cta_forest;

/* Thread-local message of the last failed call ("" if none). */
const char *cta_last_error(void);
/* Library version, major * 10000 + minor * 100 + patch. */
int cta_version(void);

/* Context on HIP device `device` with its own stream.
 * Replaces: pycuda.autoinit's implicit context (random_tree.py:1). */
int cta_ctx_create(int device, cta_ctx **out);
int cta_ctx_destroy(cta_ctx *ctx);
int cta_ctx_synchronize(cta_ctx *ctx);
/* Free / total device memory in bytes as seen by the context's device. */
int cta_ctx_mem_info(cta_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Training set: uploads X (row-major N x F, host) and compact labels y (N),
 * transposes X to feature-major on the device and presorts every feature row
 * with a stable device radix sort (ties keep ascending sample id).  idx_bits
 * 0 picks 16-bit sample indices when N <= 65,535 and 32-bit otherwise.
 * Replaces: fit_init's transpose, uploads and np.argsort
 * (random_forest.py:215-266, 233, 244-248) and get_best_dtype
 * (util.py:40-49). */
int cta_dataset_create(cta_ctx *ctx, const float *X, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                       int32_t idx_bits, cta_dataset **out);
/* As cta_dataset_create with X (N x F, float32) and y (N, uint32) in device memory. */
int cta_dataset_create_device(cta_ctx *ctx, const float *X_dev, const uint32_t *y_dev, int32_t N, int32_t F,
                              int32_t K, int32_t idx_bits, cta_dataset **out);
/* As cta_dataset_create with the class values given raw and compacted on the
 * device: keys holds N order-preserving keys of the class values (uint32
 * for key_bits 32, uint64 for 64; key order is class order, equal classes
 * have equal keys).  The distinct keys in ascending order are the classes:
 * they are written to classes (room for CTA_MAX_CLASSES uint64), their count
 * to *K, and sample i gets the index of its key as its label.  More than
 * CTA_MAX_CLASSES distinct keys fail with CTA_EUNSUPPORTED.  The label width
 * follows K (8 bits up to 255 classes, else 16) and the index width N, as in
 * cta_dataset_create.
 * Replaces: __compact_labels (random_forest.py:85-99: np.unique with
 * return_inverse on the host) and fit_init's label dtype choice (:229-231). */
int cta_dataset_create_keys(cta_ctx *ctx, const float *X, const void *keys, int32_t key_bits, int32_t N, int32_t F,
                            int32_t idx_bits, int32_t *K, uint64_t *classes, cta_dataset **out);
int cta_dataset_destroy(cta_dataset *ds);
int cta_dataset_info(const cta_dataset *ds, int32_t *N, int32_t *F, int32_t *K, int32_t *idx_bits);
/* Copies the presorted index rows (F x N, widened to uint32) to host memory.
 * Test hook for the presort. */
int cta_dataset_sorted(cta_dataset *ds, uint32_t *out);

typedef struct {
  uint64_t seed;             /* keys the per-node feature permutation */
  int32_t max_features;      /* m, 1..F (the facade applies ceil(sqrt(F)), random_forest.py:250-251) */
  int32_t min_samples_split; /* random_forest.py:48 (default 1) */
  int32_t max_depth;         /* <= 0: unlimited */
  int32_t identity_perm;     /* 1: features in id order (reference debug mode, random_tree.py:193-200) */
  int32_t mode;              /* CTA_MODE_* */
  int64_t bfs_threshold;     /* CTA_MODE_HYBRID: nodes with n > bfs_threshold go one at a time */
  int32_t scan_chunk;        /* samples per split-scan task (0: default) */
  int32_t resh_chunk;        /* samples per small-node reshuffle task (0: default) */
  int32_t trees_per_batch;   /* trees resident at once (0: as many as memory allows) */
  int32_t resh_global;       /* 1: reshuffle reads the side bitmap from global memory instead of LDS */
  int32_t bootstrap;         /* 1 with boot_ids == NULL: draw every tree's in-bag set on the device from
                                (seed, tree) -- the sorted unique values of N draws
                                floor(u * N / 2^32), u the words of Philox4x32-10((j/4, tree, "BOOT", 0), seed) */
  int32_t all_rows;          /* 1: every feature keeps presorted rows; 0 (default): two-valued features
                                other than feature 0 keep none (their one split follows from class
                                counts; the trees are identical) */
  int32_t lanes;             /* batched mode: 2 runs the batches on two streams of the context at once
                                (alternate batches, half the trees each when memory allows, each lane holding
                                at least 8 trees); 0 or 1 (default): one stream.  The trees are identical */
  int32_t fill_global;       /* 1: the side bitmaps are built with global atomics into cleared bitmaps instead
                                of in LDS (the path taken anyway when N / 8 bytes exceed a workgroup's LDS) */
  int32_t hist_layout;       /* class-count layout of the split scan: CTA_HIST_AUTO (0), CTA_HIST_REGS (1: lane k
                                holds class k, K <= 64 only) or CTA_HIST_LDS (2: per-wave LDS counters).  The
                                trees are identical */
  int32_t subtree_max;       /* batched mode: a live node of at most this many samples leaves the levels and
                                its whole subtree is grown by one wave, every such node of every tree in one
                                launch after the levels (<= 0, the default: every node stays in the levels).
                                The trees are identical */
  int32_t reserved[1];
} cta_fit_opts;

#define CTA_HIST_AUTO 0
#define CTA_HIST_REGS 1
#define CTA_HIST_LDS 2

/* Fills opts with the defaults (batched mode, m = ceil(sqrt(F)) left to the caller). */
void cta_fit_opts_default(cta_fit_opts *opts);

/* Grows T trees.  boot_ids/boot_off: tree t's in-bag set is the sorted unique
 * ids boot_ids[boot_off[t] .. boot_off[t+1]) (host memory); boot_ids == NULL
 * with opts->bootstrap == 1 draws the in-bag sets on the device; boot_ids ==
 * NULL with opts->bootstrap == 0 grows every tree on all N samples
 * (bootstrap=False).
 * Replaces: RandomForestClassifier.fit's per-tree loop (random_forest.py:292-336),
 * _get_sorted_indices with bootstrap_fill/bootstrap_reshuffle
 * (random_forest.py:113-144, bootstrap_fill.cu, bootstrap_reshuffle.cu) and
 * RandomClassifierTree.fit with its DFS/BFS kernels (random_tree.py:461-676,
 * dfs_module.cu, bfs_module.cu). */
int cta_fit(cta_ctx *ctx, cta_dataset *ds, int32_t T, const uint32_t *boot_ids, const int64_t *boot_off,
            const cta_fit_opts *opts, cta_forest **out);

int cta_forest_destroy(cta_forest *forest);
int cta_forest_info(const cta_forest *forest, int32_t *n_trees, int64_t *n_nodes, int32_t *F, int32_t *K);
/* Copies the trees out in depth-first pre-order, tree t's nodes at
 * [tree_off[t], tree_off[t+1]).  Leaves have left == right == 0, feature 0,
 * threshold 0 and value = label; internal nodes have value 0.  Any output
 * pointer except tree_off may be NULL.
 * Replaces: the tree arrays of random_tree.py:286-289, 493-513 and the pickle
 * state of random_tree.py:203-216. */
int cta_forest_export(const cta_forest *forest, int64_t *tree_off, int32_t *feature, float *threshold,
                      uint32_t *left, uint32_t *right, uint32_t *value, uint32_t *n_samples, uint32_t *depth);
/* Builds a device-resident forest from exported arrays (same layout as
 * cta_forest_export; n_samples/depth not needed).
 * Replaces: restore_forest / restore_tree (random_forest.py:17-23, random_tree.py:20-36),
 * which left the restored forest unable to predict. */
int cta_forest_import(cta_ctx *ctx, int32_t T, int32_t F, int32_t K, const int64_t *tree_off,
                      const int32_t *feature, const float *threshold, const uint32_t *left,
                      const uint32_t *right, const uint32_t *value, cta_forest **out);

/* Serialised forest: a versioned little-endian byte image of the flattened
 * node table, 48-byte header {char magic[8] = "CTAFORST", uint32 version
 * (CTA_FOREST_FORMAT), uint32 header bytes (48), int32 T, F, K, uint32 flags
 * (0), int64 nodes, uint64 0}, then tree_off (T + 1 x int64) and the node
 * arrays in the layout of cta_forest_export: feature (int32), threshold
 * (float32), left, right, value, n_samples, depth (uint32), one entry per
 * node each.  The label table (original class values) is the caller's.
 * cta_forest_save with buf == NULL reports the image size in *size.
 * cta_forest_load rejects an image of another version with
 * CTA_EUNSUPPORTED and a malformed one with CTA_EINVAL.
 * Replaces: the pickled tree arrays of random_tree.py:203-216 and the
 * restore path of random_forest.py:17-23 (which carried no format version
 * and left the restored forest unable to predict). */
#define CTA_FOREST_FORMAT 1
int cta_forest_save(const cta_forest *forest, void *buf, uint64_t cap, uint64_t *size);
int cta_forest_load(cta_ctx *ctx, const void *buf, uint64_t size, cta_forest **out);

/* Majority vote of every tree for R row-major rows (host X, R x F).  Writes
 * the winning label per row (lowest label on ties) and, if votes != NULL,
 * the per-label tree counts (R x K).
 * Replaces: predict / predict_proba (random_forest.py:339-377) with
 * BaseTree.gpu_predict and the predict kernel (base_tree.py:44-81, predict.cu). */
int cta_predict(cta_ctx *ctx, cta_forest *forest, const float *X, int64_t R, int32_t F, uint32_t *labels,
                uint32_t *votes);
/* As cta_predict with X (R x F float32), labels (R uint32) and votes (R x K
 * uint32 or NULL) in device memory. */
int cta_predict_device(cta_ctx *ctx, cta_forest *forest, const float *X_dev, int64_t R, int32_t F,
                       uint32_t *labels_dev, uint32_t *votes_dev);

/* Allocation-free predict.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  As
 * cta_predict_device without the votes output: X_dev (R x F float32) and
 * labels_dev (R uint32) in device memory; the R x K vote matrix lives in a
 * fixed per-GPU device buffer of CTA_NOALLOC_VOTE_WORDS uint32 (4 MiB, held
 * until the process ends), so R x K above that fails with CTA_EUNSUPPORTED.
 * The call allocates nothing: the forest's node table must already be on
 * the device (a fit, cta_predict or cta_predict_device puts it there, else
 * CTA_EINVAL).  The buffer keeps the last call's votes and is cleared at the
 * start of each call; calls that overlap on one GPU (any contexts) race. */
#define CTA_NOALLOC_VOTE_WORDS (1 << 20)
int cta_predict_device_noalloc(cta_ctx *ctx, cta_forest *forest, const float *X_dev, int64_t R, int32_t F,
                               uint32_t *labels_dev);

/* Frees the context's cached device workspaces (they are reused across calls
 * and grow to the largest request); *bytes reports their current size. */
int cta_ctx_release_workspace(cta_ctx *ctx);
/* Calls fn(arg) on the fitting thread after each group of levels of a
 * batched fit on this context (NULL fn: none), e.g. to let other work on the
 * device go first.  No counterpart in the reference; the serving loop uses it
 * to give queued predict batches the device between a fit's levels. */
int cta_ctx_set_level_hook(cta_ctx *ctx, void (*fn)(void *), void *arg);
int cta_ctx_workspace_bytes(cta_ctx *ctx, uint64_t *bytes);
/* Caps the context's workspaces (0: no cap).  A fit on the context then
 * sizes its batches of resident trees from what the cap leaves, and any
 * workspace growth beyond it fails with CTA_ENOMEM instead of allocating.
 * No counterpart in the reference; the serving loop holds each fit job to
 * the bytes it was admitted with. */
int cta_ctx_set_workspace_limit(cta_ctx *ctx, uint64_t bytes);
/* Device bytes of a dataset of N x F with K classes (its own buffers and the
 * transient copies of host X and y) and a bound on the context workspaces a
 * fit of T trees with these options holds (the build's sort temporaries,
 * every resident tree's rows and level state, the node arena, the canonical
 * export and the predict table).  Computed on the host; no device is needed.
 * No counterpart in the reference (which allocates per tree without
 * accounting, random_forest.py:146-174, random_tree.py:255-282); the
 * serving loop admits fit jobs with it. */
int cta_fit_estimate(int32_t N, int32_t F, int32_t K, int32_t T, const cta_fit_opts *opts, uint64_t *dataset_bytes,
                     uint64_t *workspace_bytes);

/* ---- serving loop -------------------------------------------------------
 * No reference counterpart: the reference re-uploads rows and trees on every
 * predict call (base_tree.py:57-62).  A server keeps forests resident, runs a
 * predict worker and fit workers on their own contexts and streams, coalesces
 * the rows of queued predict requests for one forest into one traversal
 * launch (requests arriving while a batch runs join the next batch), and
 * admits fit jobs only when their estimated device footprint fits the HBM
 * budget (a job larger than the whole budget fails with CTA_ENOMEM). */
typedef struct cta_server cta_server;

typedef struct {
  int32_t device;
  int32_t coalesce;        /* 1: one launch for the queued requests of a forest; 0: one launch per request */
  int64_t max_batch_rows;  /* rows per launch at most (0: 1,048,576); also the largest request */
  uint64_t hbm_budget;     /* device bytes the server may hold (0: 90% of free memory at creation) */
  int32_t fit_workers;     /* concurrent fit jobs (0: 1) */
  int32_t pad0;
  int64_t max_batch_work;  /* coalesced rows x trees per launch at most (0: 2^22, about 1 ms of
                              traversal); a single request may exceed it */
  int32_t fit_yield_us;    /* a fit job waits up to this long after each group of levels while
                              predict work is queued or running (0: 4,000; negative: never) */
  int32_t split_rows;      /* with coalescing, requests of at most this many rows go to a short-request
                              lane and larger ones to a long-request lane of their own (two predict
                              workers, contexts and streams), so short requests never wait behind a long
                              traversal (0: 4,096; negative: one lane) */
  int32_t reserved[2];
} cta_server_opts;

typedef struct {
  uint64_t predict_requests, predict_rows, predict_launches;
  uint64_t fit_jobs, fit_rejected;
  double p50_ms, p99_ms, mean_ms;  /* predict latency, submission to completion */
  double rows_per_s;               /* predict rows over first submission .. last completion */
  uint64_t hbm_budget, hbm_in_use, hbm_peak;
  int32_t resident_forests, queued_predict, queued_fit, reserved;
  uint64_t fit_mem_measured_max;   /* largest drop of the device's free memory over a fit job (dataset build
                                      to the end of its fit) */
  double fit_mem_ratio_max;        /* largest such drop over the bytes the job was admitted with */
} cta_server_stats;

void cta_server_opts_default(cta_server_opts *opts);
/* Thread-local message of the last failed cta_server_* call. */
const char *cta_server_last_error(void);
int cta_server_create(const cta_server_opts *opts, cta_server **out);
/* Drains queued work, stops the workers and frees every resident forest. */
int cta_server_destroy(cta_server *server);
/* Makes a resident copy of `forest` on the server's device. */
int cta_server_add_forest(cta_server *server, const cta_forest *forest, int32_t *forest_id);
/* Borrowed handle of a resident forest (valid until it is dropped), e.g. to
 * export a forest fitted by the server. */
int cta_server_get_forest(cta_server *server, int32_t forest_id, cta_forest **forest);
/* Frees a resident forest once its queued requests are done. */
int cta_server_drop_forest(cta_server *server, int32_t forest_id);
/* Queues R rows (host, R x F, valid until the ticket completes); the labels
 * (R uint32) are written before the ticket completes. */
int cta_server_submit_predict(cta_server *server, int32_t forest_id, const float *X, int64_t R, int32_t F,
                              uint32_t *labels, uint64_t *ticket);
/* Queues `count` requests at once (request i: R[i] rows at X[i], labels to
 * labels[i], ticket to tickets[i]), as count calls of
 * cta_server_submit_predict under one lock; all are refused if any is bad. */
int cta_server_submit_predict_many(cta_server *server, int32_t forest_id, int32_t count, const float *const *X,
                                   const int64_t *R, int32_t F, uint32_t *const *labels, uint64_t *tickets);
/* Queues a fit job (X, y copied at submission; boot_ids come from
 * opts->bootstrap).  On success the new forest becomes resident and its id is
 * reported by cta_server_wait. */
int cta_server_submit_fit(cta_server *server, const float *X, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                          int32_t T, const cta_fit_opts *opts, uint64_t *ticket);
/* Waits up to timeout_us (< 0: forever).  *status: -1 still pending, else the
 * job's CTA_* status; *forest_id: the fitted forest of a fit job.  A
 * completed ticket is forgotten after it has been reported. */
int cta_server_wait(cta_server *server, uint64_t ticket, int64_t timeout_us, int32_t *status, int32_t *forest_id);
/* cta_server_wait over `count` tickets in order, with one deadline
 * timeout_us from the call (< 0: forever): status[i] (and forest_ids[i] when
 * forest_ids is not NULL) as cta_server_wait reports them; tickets still
 * pending at the deadline report -1 and are kept.  All are refused if any
 * ticket is unknown. */
int cta_server_wait_many(cta_server *server, int32_t count, const uint64_t *tickets, int64_t timeout_us,
                         int32_t *status, int32_t *forest_ids);
int cta_server_get_stats(cta_server *server, cta_server_stats *out);
int cta_server_reset_stats(cta_server *server);

/* ---- measurement ------------------------------------------------------- */
typedef struct {
  char name[32];
  uint64_t launches;      /* timed launches */
  double total_ms;        /* sum of HIP-event durations on the context stream */
  uint64_t alg_bytes;     /* algorithmic bytes of the timed launches */
} cta_kernel_stat;

/* Enables per-launch HIP-event timing of the fit and predict kernels. */
int cta_ctx_set_profiling(cta_ctx *ctx, int enable);
int cta_ctx_reset_stats(cta_ctx *ctx);
/* Copies up to max entries; *n gets the number of kernels with stats. */
int cta_ctx_kernel_stats(cta_ctx *ctx, cta_kernel_stat *out, int32_t max, int32_t *n);

/* ---- per-kernel parity hooks (tests) ------------------------------------ */
/* Split scan of one node: rows (F x n uint32, host) hold the node's sorted
 * segment of every feature; the m_eval features in `features` are scanned
 * with the device kernel.  Writes the winning feature (-1 if none), position
 * and exact score numerator/denominator.
 * Replaces: scan_gini_large / scan_reduce / compute_2d / reduce_2d /
 * find_min_imp (dfs_module.cu:24-208) and scan_bfs / compute_2d / reduce
 * (bfs_module.cu:25-235) for one node. */
int cta_k_split_scan(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const int32_t *features,
                     int32_t m_eval, int32_t chunk, int32_t *best_feature, uint32_t *best_pos, uint64_t *num,
                     uint64_t *den);
/* Stable partition of every row of one node (rows F x n uint32, host) by
 * left-membership of each sample (mark, N bytes, 1 = left) into out (F x n).
 * chunk: samples per small-node task (0: default); global_bits: 1 reads the
 * side bitmap from global memory instead of staging it in LDS.
 * Replaces: fill_table + scan_reshuffle (dfs_module.cu:210-307, bfs_module.cu:238-377). */
int cta_k_reshuffle(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const uint8_t *mark,
                    int32_t chunk, int32_t global_bits, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif

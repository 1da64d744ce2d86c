/*
 * oracle.h -- CPU restatement of CudaTree's random-forest fit/predict path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in cudatree_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * It restates, single-threaded and in plain C, the canonical semantics of the
 * reference algorithm (cudatree/random_forest.py, cudatree/random_tree.py and
 * cudatree/cuda_kernels/{dfs,bfs}_module.cu, predict.cu, bootstrap_*.cu).
 * Every function cites the reference lines it follows.  Where the reference's
 * behaviour depends on unpinnable state (numpy's global RNG, unstable argsort,
 * float-rounded tie order), the canonical rule used here is stated at the
 * function.
 *
 * Pinning status (which independent check fixes each function):
 *   orc_philox4x32_10   -- Random123 known-answer vectors (tests/test_oracle.py)
 *   orc_feature_perm    -- permutation/identity/determinism invariants
 *   orc_bootstrap_ids   -- sorted/unique/range invariants, expected in-bag
 *                          fraction 1 - (1 - 1/N)^N, uniform coverage
 *   orc_presort         -- numpy.argsort(kind="stable") (closed form)
 *   orc_scan_feature    -- brute-force exact-fraction Gini on tiny inputs
 *   orc_fit             -- brute-force tree builder on tiny inputs; memorize
 *                          property on sklearn-bundled datasets (the reference's
 *                          own tests); accuracy within 2% of sklearn
 *   orc_predict         -- brute-force traversal + vote on random forests
 * No function is left "parity unpinned".
 */
#ifndef CUDATREE_ORACLE_H
#define CUDATREE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with the product C-ABI. */
#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2

/* Largest sample count whose Gini numerators fit the 64-bit exact arithmetic
 * (num <= n^3/4 < 2^64 for n < 2^22). */
#define ORC_MAX_SAMPLES 4194303

typedef struct {
  uint64_t seed;           /* keys the per-node feature permutation */
  int32_t max_features;    /* m: non-constant features evaluated per node */
  int32_t min_samples_split;
  int32_t max_depth;       /* <= 0: unlimited (the reference has no cap) */
  int32_t identity_perm;   /* 1: reference debug mode, features in id order */
} orc_params;

/* A fitted forest: T trees, nodes of tree t at [tree_off[t], tree_off[t+1]),
 * numbered in depth-first pre-order (root first, left subtree, right subtree).
 * Leaves have left == right == 0, feature 0, threshold 0 and value = label.
 * Internal nodes have value 0.  n_samples is the node's sample count. */
typedef struct {
  int32_t n_trees;
  int64_t *tree_off;
  int32_t *feature;
  float *threshold;
  uint32_t *left;
  uint32_t *right;
  uint32_t *value;
  uint32_t *n_samples;
  uint32_t *depth;
} orc_forest;

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Canonical per-node feature permutation pi over [0, F) keyed by
 * (seed, tree, depth, start); identity when identity != 0. */
void orc_feature_perm(uint64_t seed, uint32_t tree, uint32_t depth, uint32_t start,
                      int32_t F, int32_t identity, int32_t *out);

/* Canonical bootstrap of tree `tree`: the sorted unique values of N draws
 * u_j -> floor(u_j * N / 2^32), u_j the (j mod 4)-th word of
 * Philox4x32-10(counter = (j / 4, tree, "BOOT", 0), key = seed).  Writes the
 * ids to out (capacity N) and returns their count. */
int64_t orc_bootstrap_ids(uint64_t seed, uint32_t tree, int32_t N, uint32_t *out);

/* Stable ascending argsort of every feature row of X_T (F x N, feature-major).
 * Ties (including -0.0 == +0.0) keep ascending sample id.  NaN -> ORC_EINVAL. */
int orc_presort(const float *X_T, int32_t N, int32_t F, uint32_t *out);

/* Gini scan of one feature over a node's sorted segment (values v[0..n),
 * labels lab[0..n) in [0,K)).  Returns 1 and the best position p (left child =
 * [0..p]) with its exact score Q = num/den, or 0 if no position is eligible. */
int orc_scan_feature(const float *v, const uint32_t *lab, uint32_t n, uint32_t K,
                     uint32_t *p, uint64_t *num, uint64_t *den);

/* Fit T trees.  sorted: F x N presorted indices (NULL: computed here).
 * boot_ids/boot_off: sorted unique in-bag ids of tree t at
 * boot_ids[boot_off[t] .. boot_off[t+1]); boot_ids == NULL means every tree
 * uses all N samples (bootstrap=False). */
int orc_fit(const float *X_T, const uint32_t *y, int32_t N, int32_t F, int32_t K,
            const uint32_t *sorted, int32_t T, const uint32_t *boot_ids,
            const int64_t *boot_off, const orc_params *prm, orc_forest *out);
/* As orc_fit with the trees grown by `threads` host threads at once (each
 * tree is independent; the forest is the same, trees in order). */
int orc_fit_threads(const float *X_T, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                    const uint32_t *sorted, int32_t T, const uint32_t *boot_ids,
                    const int64_t *boot_off, const orc_params *prm, int32_t threads,
                    orc_forest *out);

void orc_forest_free(orc_forest *f);

/* Majority vote over the trees for R row-major rows X (R x F).  out_labels[r]
 * gets the winning label (lowest on ties); votes (R x K, may be NULL) gets
 * the per-label tree counts. */
int orc_predict(const orc_forest *f, const float *X, int64_t R, int32_t F, int32_t K,
                uint32_t *out_labels, uint32_t *votes);

#ifdef __cplusplus
}
#endif
#endif

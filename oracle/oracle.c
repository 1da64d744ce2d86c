/*
 * oracle.c -- CPU restatement of CudaTree's random-forest path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Loaded
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product path.  Single-threaded, written for clarity over speed.
 *
 * Canonical semantics restated here (reference file:line in brackets):
 *   D1 labels are compacted to 0..K-1 by the caller      [random_forest.py:85-99]
 *   D3 bootstrap = sorted unique in-bag ids, fed in      [random_forest.py:113-144]
 *   D4 presort is a stable argsort per feature row       [random_forest.py:247]
 *   D5 position p is a candidate iff
 *        (double)s[p] + 1e-5 < (double)s[p+1]           [common_func.cu:1, bfs_module.cu:173]
 *   D6 threshold = (s[p] + s[p+1]) * 0.5f in float       [bfs_module.cu:399-400]
 *   D7 leaf label = label of the first sample of the leaf's segment in
 *        feature 0's sorted row                          [random_tree.py:50-68, 493-513]
 *   D8 features are visited in a per-node permutation; constant features
 *        (first == last value in the node) are skipped until m have been
 *        evaluated or the permutation is exhausted       [bfs_module.cu:122-139]
 *   D9 the split maximises Q = S_L/n_L + S_R/n_R (S = sum of squared class
 *        counts), i.e. minimises the weighted Gini of   [dfs_module.cu:384-428]
 *        the reference, compared exactly on integers; ties go to the lowest
 *        feature id, then the lowest position
 *   D10 a node is a leaf if pure, if n < min_samples_split, at max_depth, or
 *        if no candidate exists                          [random_tree.py:587-634, 113-163]
 *   D11 every feature row is stably partitioned into [start, start+p+1) and
 *        [start+p+1, stop)                               [dfs_module.cu:224-307, bfs_module.cu:266-377]
 *   D12 predict: go left iff x[f] < threshold; the vote is argmax of the
 *        per-label tree count, lowest label on ties      [predict.cu:50-67, random_forest.py:339-362]
 *
 * Compile with -DORC_MUTANT=k (k >= 1) to plant a deliberate mistake; the
 * test suite uses those builds to show that its checks catch each one.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <stdint.h>

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, from Salmon et al., "Parallel random numbers: as easy as   */
/* 1, 2, 3" (SC'11): 10 rounds of two 32x32->64 multiplies with a Weyl key  */
/* schedule.  The reference draws its feature shuffle from numpy's global   */
/* generator (random_tree.py:46-48); a counter-based generator makes the    */
/* permutation a pure function of the node, independent of build order.     */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static void philox_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
  uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k[0] += PHILOX_W0;
      k[1] += PHILOX_W1;
    }
    philox_round(c, k);
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = c[3];
}

/* ------------------------------------------------------------------------ */
/* Feature permutation (canonical D8).                                      */
/* The reference keeps one cumulative Sattolo-style shuffle of the feature  */
/* ids (random_tree.py:38-48, 198-200) re-drawn after every node or level.  */
/* Here pi is a 4-round Feistel network over 2h bits (2^(2h) >= F) with     */
/* round keys from one Philox block keyed by (seed, tree, depth, start),   */
/* restricted to [0, F) by cycle walking.  identity != 0 reproduces the     */
/* reference's debug mode (features 0..F-1, random_tree.py:193-200).        */
/* ------------------------------------------------------------------------ */
#define PERM_TAG 0x5045524Du /* "PERM" */

static uint32_t fmix32(uint32_t x) { /* MurmurHash3 finalizer (Appleby) */
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

static uint32_t feistel(uint32_t x, uint32_t h, const uint32_t rk[4]) {
  uint32_t mask = (1u << h) - 1u;
  uint32_t L = x >> h, R = x & mask;
  for (int r = 0; r < 4; ++r) {
    uint32_t f = fmix32(R ^ rk[r]) & mask;
    uint32_t nl = R;
    R = L ^ f;
    L = nl;
  }
  return (L << h) | R;
}

void orc_feature_perm(uint64_t seed, uint32_t tree, uint32_t depth, uint32_t start,
                      int32_t F, int32_t identity, int32_t *out) {
  if (identity) {
    for (int32_t k = 0; k < F; ++k) out[k] = k;
    return;
  }
  uint32_t h = 1;
  while ((1ull << (2 * h)) < (uint64_t)F) ++h;
  uint32_t ctr[4] = {start, depth, tree, PERM_TAG};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t rk[4];
  orc_philox4x32_10(ctr, key, rk);
  for (int32_t k = 0; k < F; ++k) {
    uint32_t y = feistel((uint32_t)k, h, rk);
    while (y >= (uint32_t)F) y = feistel(y, h, rk);
    out[k] = (int32_t)y;
  }
}

/* ------------------------------------------------------------------------ */
/* Bootstrap (canonical D3).  The reference draws np.unique(np.random.randint */
/* (0, N, N)) from numpy's global generator (random_forest.py:123-126): the  */
/* set of distinct ids among N uniform draws.  Here the draws come from      */
/* Philox keyed by (seed, tree), four per block, mapped to [0, N) by the     */
/* high half of u * N.                                                       */
/* ------------------------------------------------------------------------ */
#define BOOT_TAG 0x424F4F54u /* "BOOT" */

int64_t orc_bootstrap_ids(uint64_t seed, uint32_t tree, int32_t N, uint32_t *out) {
  if (N <= 0) return 0;
  uint8_t *seen = (uint8_t *)calloc((size_t)N, 1);
  if (!seen) return -1;
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  for (int64_t j = 0; j < N; ++j) {
    if ((j & 3) == 0) {
      uint32_t ctr[4] = {(uint32_t)(j >> 2), tree, BOOT_TAG, 0u};
      orc_philox4x32_10(ctr, key, w);
    }
    uint32_t id = (uint32_t)(((uint64_t)w[j & 3] * (uint64_t)N) >> 32);
    seen[id] = 1;
  }
  int64_t n = 0;
  for (int32_t i = 0; i < N; ++i)
    if (seen[i]) out[n++] = (uint32_t)i;
  free(seen);
  return n;
}

/* ------------------------------------------------------------------------ */
/* Presort (canonical D4).  The reference sorts each feature row of the     */
/* transposed samples with numpy's default argsort (random_forest.py:233,  */
/* 247); ties there are unordered.  Here ties keep ascending sample id.     */
/* ------------------------------------------------------------------------ */
typedef struct {
  float v;
  uint32_t i;
} keyed;

static int keyed_cmp(const void *a, const void *b) {
  const keyed *x = (const keyed *)a, *y = (const keyed *)b;
  if (x->v < y->v) return -1;
  if (x->v > y->v) return 1;
  return (x->i < y->i) ? -1 : (x->i > y->i);
}

int orc_presort(const float *X_T, int32_t N, int32_t F, uint32_t *out) {
  if (N <= 0 || F <= 0) return ORC_EINVAL;
  keyed *buf = (keyed *)malloc(sizeof(keyed) * (size_t)N);
  if (!buf) return ORC_ENOMEM;
  for (int32_t f = 0; f < F; ++f) {
    const float *row = X_T + (size_t)f * N;
    for (int32_t i = 0; i < N; ++i) {
      if (isnan(row[i])) {
        free(buf);
        return ORC_EINVAL;
      }
      buf[i].v = row[i];
      buf[i].i = (uint32_t)i;
    }
    /* keyed_cmp is a total order (no NaN), so any sort is the stable one. */
    qsort(buf, (size_t)N, sizeof(keyed), keyed_cmp);
    for (int32_t i = 0; i < N; ++i) out[(size_t)f * N + i] = buf[i].i;
  }
  free(buf);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Exact score comparison (canonical D9).                                   */
/* The reference scores a position by the weighted Gini                     */
/*   G = (1 - S_L/n_L^2) n_L/n + (1 - S_R/n_R^2) n_R/n = 1 - Q/n            */
/* in mixed float/double and keeps strict minima (dfs_module.cu:404-428,   */
/* bfs_module.cu:176-187).  Minimising G is maximising                      */
/*   Q = S_L/n_L + S_R/n_R = (S_L n_R + S_R n_L) / (n_L n_R) = num/den,     */
/* compared here by 128-bit cross multiplication.                           */
/* ------------------------------------------------------------------------ */
static int q_greater(uint64_t n1, uint64_t d1, uint64_t n2, uint64_t d2) {
  unsigned __int128 a = (unsigned __int128)n1 * d2;
  unsigned __int128 b = (unsigned __int128)n2 * d1;
  return a > b;
}

static int q_equal(uint64_t n1, uint64_t d1, uint64_t n2, uint64_t d2) {
  unsigned __int128 a = (unsigned __int128)n1 * d2;
  unsigned __int128 b = (unsigned __int128)n2 * d1;
  return a == b;
}

/* D5: the reference skips position t when
 *   shared_samples[t] + DIFF_THRESHOLD >= shared_samples[t + 1]
 * with DIFF_THRESHOLD the double literal 0.00001 (common_func.cu:1,
 * dfs_module.cu:400, bfs_module.cu:173): float operands promoted to double. */
static int candidate(float a, float b) {
#if ORC_MUTANT == 3
  return a + 0.00001f < b;
#else
  return (double)a + 0.00001 < (double)b;
#endif
}

/* Gini scan over one feature's sorted segment, following the serial loop of
 * compute_2d (dfs_module.cu:384-428, bfs_module.cu:160-188): the left sum of
 * squared counts grows by 2c+1 when a sample of a class with c samples
 * already on the left moves over (dfs_module.cu:396-397); the right sum is
 * kept as S_R = T2 - 2 D + S_L with T2 = sum tot_k^2 and D = sum tot_k c_k. */
static int scan_segment(const float *v, const uint32_t *lab, uint32_t n, const uint32_t *tot,
                        uint32_t K, uint32_t *cnt, uint32_t *bp, uint64_t *bnum, uint64_t *bden) {
  uint64_t T2 = 0, SL = 0, D = 0;
  int found = 0;
  for (uint32_t k = 0; k < K; ++k) {
    T2 += (uint64_t)tot[k] * tot[k];
    cnt[k] = 0;
  }
  for (uint32_t p = 0; p + 1 < n; ++p) {
    uint32_t l = lab[p];
#if ORC_MUTANT == 1
    SL += 2ull * cnt[l];
#else
    SL += 2ull * cnt[l] + 1ull;
#endif
    D += tot[l];
    cnt[l] += 1;
    if (!candidate(v[p], v[p + 1])) continue;
#if ORC_MUTANT == 8
    uint64_t nL = p, nR = (uint64_t)n - p;
    if (nL == 0) continue;
#else
    uint64_t nL = (uint64_t)p + 1, nR = (uint64_t)n - nL;
#endif
#if ORC_MUTANT == 2
    uint64_t SR = T2 + 2 * D - SL;
#else
    uint64_t SR = T2 + SL - 2 * D;
#endif
    uint64_t num = SL * nR + SR * nL;
    uint64_t den = nL * nR;
    if (!found || q_greater(num, den, *bnum, *bden)) {
      found = 1;
      *bp = p;
      *bnum = num;
      *bden = den;
    }
  }
  return found;
}

int orc_scan_feature(const float *v, const uint32_t *lab, uint32_t n, uint32_t K, uint32_t *p,
                     uint64_t *num, uint64_t *den) {
  if (n == 0 || K == 0 || n > ORC_MAX_SAMPLES) return 0;
  uint32_t *tot = (uint32_t *)calloc(K, sizeof(uint32_t));
  uint32_t *cnt = (uint32_t *)calloc(K, sizeof(uint32_t));
  if (!tot || !cnt) {
    free(tot);
    free(cnt);
    return 0;
  }
  for (uint32_t i = 0; i < n; ++i) tot[lab[i]]++;
  int r = scan_segment(v, lab, n, tot, K, cnt, p, num, den);
  free(tot);
  free(cnt);
  return r;
}

/* ------------------------------------------------------------------------ */
/* Tree construction                                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  /* inputs */
  const float *X_T;
  const uint32_t *y;
  int32_t N, F, K;
  const orc_params *prm;
  uint32_t tree;
  /* ping-pong sorted-index rows, stride n_root (random_tree.py:466-467) */
  uint32_t *buf[2];
  uint32_t stride;
  uint8_t *mark; /* mark table, one byte per sample (random_forest.py:163) */
  /* scratch */
  int32_t *perm;
  float *vals;
  uint32_t *labs;
  uint32_t *cnt;
  /* growing node arrays */
  int64_t n_nodes, cap;
  int32_t *feature;
  float *threshold;
  uint32_t *left, *right, *value, *n_samples, *depth;
} tree_ctx;

typedef struct {
  uint32_t start, n, depth, parity;
  int64_t parent; /* -1 for the root */
  int32_t is_right;
} work;

static int grow(tree_ctx *t) {
  if (t->n_nodes < t->cap) return 0;
  int64_t nc = t->cap ? t->cap * 2 : 64;
#define GROW(p, T_)                                     \
  do {                                                  \
    void *q = realloc(t->p, sizeof(T_) * (size_t)nc);   \
    if (!q) return -1;                                  \
    t->p = (T_ *)q;                                     \
  } while (0)
  GROW(feature, int32_t);
  GROW(threshold, float);
  GROW(left, uint32_t);
  GROW(right, uint32_t);
  GROW(value, uint32_t);
  GROW(n_samples, uint32_t);
  GROW(depth, uint32_t);
#undef GROW
  t->cap = nc;
  return 0;
}

/* Leaf label of the first sample of the segment in feature 0's row (D7):
 * decorate() reads target[si[values_idx]] with values_idx the segment start
 * (random_tree.py:50-63, 65-68). */
static uint32_t first_sample_label(const tree_ctx *t, const uint32_t *rows, uint32_t start, uint32_t n) {
#if ORC_MUTANT == 7
  return t->y[rows[start + n - 1]];
#else
  (void)n;
  return t->y[rows[start]];
#endif
}

static uint32_t majority(const uint32_t *counts, uint32_t K) {
  uint32_t best = 0;
  for (uint32_t k = 1; k < K; ++k)
    if (counts[k] > counts[best]) best = k;
  return best;
}

/* Stable partition of one row segment by the mark table: marked (left)
 * samples keep their order in [start, start+nL), the rest in
 * [start+nL, start+n) (dfs_module.cu:299-300, bfs_module.cu:364-366). */
static void partition_row(const uint32_t *in, uint32_t *out, const uint8_t *mark, uint32_t start,
                          uint32_t n, uint32_t nL) {
  uint32_t l = start, r = start + nL;
#if ORC_MUTANT == 6
  r = start + n;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t s = in[start + i];
    if (mark[s]) out[l++] = s;
    else out[--r] = s;
  }
#else
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t s = in[start + i];
    if (mark[s]) out[l++] = s;
    else out[r++] = s;
  }
#endif
}

/* One node: leaf tests, feature walk, scans, split and partition.  This is
 * the work of __dfs_construct (random_tree.py:580-676) and of one node's
 * share of __bfs (random_tree.py:313-458) under the canonical rules above.
 * Returns 1 if the node was split (children to be built), 0 for a leaf,
 * -1 on allocation failure. */
static int build_node(tree_ctx *t, const work *w, const uint32_t *counts, int64_t id, uint32_t *cntL,
                      uint32_t *out_nL) {
  const int32_t F = t->F, K = t->K;
  const uint32_t n = w->n, start = w->start;
  const uint32_t *cur = t->buf[w->parity];
  uint32_t *nxt = t->buf[w->parity ^ 1];
  const orc_params *prm = t->prm;

  t->feature[id] = 0;
  t->threshold[id] = 0.0f;
  t->left[id] = 0;
  t->right[id] = 0;
  t->value[id] = 0;
  t->n_samples[id] = n;
  t->depth[id] = w->depth;

  /* D10: pure node -> leaf with its class (the reference checks the child
   * impurity reported by the parent, random_tree.py:587-605, 123-163). */
  uint32_t nz = 0, only = 0;
  for (int32_t k = 0; k < K; ++k)
    if (counts[k]) {
      ++nz;
      only = (uint32_t)k;
    }
  if (nz <= 1) {
    t->value[id] = only;
    return 0;
  }
  if ((int64_t)n < (int64_t)prm->min_samples_split) { /* random_tree.py:607-614 */
    t->value[id] = first_sample_label(t, cur, start, n);
    return 0;
  }
  if (prm->max_depth > 0 && (int64_t)w->depth >= (int64_t)prm->max_depth) {
    t->value[id] = majority(counts, (uint32_t)K);
    return 0;
  }

  /* D8: walk the permutation, skip features constant in the node
   * (bfs_module.cu:130-139), evaluate until m have been scanned. */
  orc_feature_perm(prm->seed, t->tree, w->depth, start, F, prm->identity_perm, t->perm);
  int32_t evaluated = 0, bf = -1;
  uint32_t bp = 0;
  uint64_t bnum = 0, bden = 1;
  for (int32_t k = 0; k < F && evaluated < prm->max_features; ++k) {
    int32_t f = t->perm[k];
    const uint32_t *row = cur + (size_t)f * t->stride + start;
    const float *x = t->X_T + (size_t)f * t->N;
#if ORC_MUTANT != 9
    if (x[row[0]] == x[row[n - 1]]) continue;
#endif
    ++evaluated;
    for (uint32_t i = 0; i < n; ++i) {
      t->vals[i] = x[row[i]];
      t->labs[i] = t->y[row[i]];
    }
    uint32_t p;
    uint64_t num, den;
    if (!scan_segment(t->vals, t->labs, n, counts, (uint32_t)K, t->cnt, &p, &num, &den)) continue;
    int better = bf < 0 || q_greater(num, den, bnum, bden);
#if ORC_MUTANT == 5
    if (!better && q_equal(num, den, bnum, bden) && f > bf) better = 1;
#else
    if (!better && q_equal(num, den, bnum, bden) && f < bf) better = 1;
#endif
    if (better) {
      bf = f;
      bp = p;
      bnum = num;
      bden = den;
    }
  }
  if (bf < 0) { /* no candidate: random_tree.py:628-634, 113-114 */
    t->value[id] = first_sample_label(t, cur, start, n);
    return 0;
  }

  /* D6: threshold midway between the two sorted neighbours, in float
   * (get_thresholds, bfs_module.cu:399-400). */
  const uint32_t *brow = cur + (size_t)bf * t->stride + start;
  const float *bx = t->X_T + (size_t)bf * t->N;
  float a = bx[brow[bp]], b = bx[brow[bp + 1]];
#if ORC_MUTANT == 4
  float thr = a;
  (void)b;
#else
  volatile float sum = a + b; /* no contraction: round the sum, then halve */
  float thr = sum * 0.5f;
#endif
  t->feature[id] = bf;
  t->threshold[id] = thr;

  /* Mark table (fill_table, dfs_module.cu:210-221, bfs_module.cu:238-264)
   * and the left child's class counts. */
  for (int32_t k = 0; k < K; ++k) cntL[k] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t s = brow[i];
    t->mark[s] = (uint8_t)(i <= bp);
    if (i <= bp) cntL[t->y[s]]++;
  }
  /* D11: every feature row, into the other ping-pong buffer. */
  for (int32_t f = 0; f < F; ++f)
    partition_row(cur + (size_t)f * t->stride, nxt + (size_t)f * t->stride, t->mark, start, n, bp + 1);
  *out_nL = bp + 1;
  return 1;
}

static void free_tree_ctx(tree_ctx *t) {
  free(t->buf[0]);
  free(t->buf[1]);
  free(t->mark);
  free(t->perm);
  free(t->vals);
  free(t->labs);
  free(t->cnt);
  free(t->feature);
  free(t->threshold);
  free(t->left);
  free(t->right);
  free(t->value);
  free(t->n_samples);
  free(t->depth);
}

/* Build one tree in depth-first pre-order with an explicit stack (the
 * reference recurses, random_tree.py:670-676).  Stack entry i owns the
 * class-count slot i; popping entry i and pushing its children at i (right)
 * and i+1 (left) keeps every waiting right sibling's counts intact.
 * Appends the tree to *out. */
static int build_tree(tree_ctx *t, const uint32_t *sorted, const uint32_t *ids, int64_t n_ids, orc_forest *out) {
  const int32_t F = t->F, K = t->K, N = t->N;
  uint32_t n_root = (uint32_t)(ids ? n_ids : N);
  int rc = ORC_OK;
  t->stride = n_root;
  t->buf[0] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)F * n_root);
  t->buf[1] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)F * n_root);
  t->mark = (uint8_t *)calloc((size_t)N, 1);
  t->perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)F);
  t->vals = (float *)malloc(sizeof(float) * n_root);
  t->labs = (uint32_t *)malloc(sizeof(uint32_t) * n_root);
  t->cnt = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)K);
  int64_t scap = 64;
  work *stack = (work *)malloc(sizeof(work) * (size_t)scap);
  uint32_t *slots = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)K * (size_t)scap);
  uint32_t *wc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)K);
  if (!t->buf[0] || !t->buf[1] || !t->mark || !t->perm || !t->vals || !t->labs || !t->cnt || !stack || !slots ||
      !wc) {
    rc = ORC_ENOMEM;
    goto done;
  }

  /* Bootstrap compaction (bootstrap_fill.cu:6-19, bootstrap_reshuffle.cu:9-59):
   * mark the in-bag ids, keep them in presorted order in every row. */
  if (ids) {
    memset(t->mark, 0, (size_t)N);
    for (int64_t j = 0; j < n_ids; ++j) t->mark[ids[j]] = 1;
    for (int32_t f = 0; f < F; ++f) {
      uint32_t o = 0;
      const uint32_t *src = sorted + (size_t)f * N;
      for (int32_t i = 0; i < N; ++i)
        if (t->mark[src[i]]) t->buf[0][(size_t)f * n_root + o++] = src[i];
    }
  } else {
    memcpy(t->buf[0], sorted, sizeof(uint32_t) * (size_t)F * N);
  }

  int64_t sp = 0;
  for (int32_t k = 0; k < K; ++k) slots[k] = 0;
  for (uint32_t i = 0; i < n_root; ++i) slots[t->y[t->buf[0][i]]]++;
  stack[sp++] = (work){0, n_root, 0, 0, -1, 0};
  while (sp > 0) {
    work w = stack[--sp];
    memcpy(wc, slots + (size_t)K * (size_t)sp, sizeof(uint32_t) * (size_t)K);
    if (sp + 2 > scap) {
      int64_t nc = scap * 2;
      work *s2 = (work *)realloc(stack, sizeof(work) * (size_t)nc);
      if (!s2) {
        rc = ORC_ENOMEM;
        goto done;
      }
      stack = s2;
      uint32_t *c2 = (uint32_t *)realloc(slots, sizeof(uint32_t) * (size_t)K * (size_t)nc);
      if (!c2) {
        rc = ORC_ENOMEM;
        goto done;
      }
      slots = c2;
      scap = nc;
    }
    if (grow(t)) {
      rc = ORC_ENOMEM;
      goto done;
    }
    int64_t id = t->n_nodes++;
    if (w.parent >= 0) {
      if (w.is_right) t->right[w.parent] = (uint32_t)id;
      else t->left[w.parent] = (uint32_t)id;
    }
    uint32_t *cr = slots + (size_t)K * (size_t)sp;
    uint32_t *cl = slots + (size_t)K * (size_t)(sp + 1);
    uint32_t nL = 0;
    int r = build_node(t, &w, wc, id, cl, &nL);
    if (r == 1) {
      for (int32_t k = 0; k < K; ++k) cr[k] = wc[k] - cl[k];
      /* right first, so the left subtree is numbered next (pre-order) */
      stack[sp] = (work){w.start + nL, w.n - nL, w.depth + 1, w.parity ^ 1, id, 1};
      stack[sp + 1] = (work){w.start, nL, w.depth + 1, w.parity ^ 1, id, 0};
      sp += 2;
    }
  }

  {
    int32_t tt = out->n_trees;
    int64_t base = out->tree_off[tt];
    int64_t total = base + t->n_nodes;
#define APPEND(field, T_)                                                      \
  do {                                                                         \
    void *q = realloc(out->field, sizeof(T_) * (size_t)(total ? total : 1));   \
    if (!q) {                                                                  \
      rc = ORC_ENOMEM;                                                         \
      goto done;                                                               \
    }                                                                          \
    out->field = (T_ *)q;                                                      \
    memcpy(out->field + base, t->field, sizeof(T_) * (size_t)t->n_nodes);      \
  } while (0)
    APPEND(feature, int32_t);
    APPEND(threshold, float);
    APPEND(left, uint32_t);
    APPEND(right, uint32_t);
    APPEND(value, uint32_t);
    APPEND(n_samples, uint32_t);
    APPEND(depth, uint32_t);
#undef APPEND
    out->tree_off[tt + 1] = total;
    out->n_trees = tt + 1;
  }
done:
  free(stack);
  free(slots);
  free(wc);
  return rc;
}

/* One tree (tree index t) appended to `out`. */
static int fit_one_tree(const float *X_T, const uint32_t *y, int32_t N, int32_t F, int32_t K, const orc_params *prm,
                        const uint32_t *sorted, int32_t t, const uint32_t *boot_ids, const int64_t *boot_off,
                        orc_forest *out) {
  tree_ctx tc;
  memset(&tc, 0, sizeof(tc));
  tc.X_T = X_T;
  tc.y = y;
  tc.N = N;
  tc.F = F;
  tc.K = K;
  tc.prm = prm;
  tc.tree = (uint32_t)t;
  const uint32_t *ids = boot_ids ? boot_ids + boot_off[t] : NULL;
  int64_t n_ids = boot_ids ? boot_off[t + 1] - boot_off[t] : 0;
  int rc = build_tree(&tc, sorted, ids, n_ids, out);
  free_tree_ctx(&tc);
  return rc;
}

/* Appends every tree of `src` to `dst`. */
static int append_forest(orc_forest *dst, const orc_forest *src) {
  const int64_t base = dst->tree_off[dst->n_trees], n = src->tree_off[src->n_trees], total = base + n;
#define APPEND2(field, T_)                                                     \
  do {                                                                         \
    void *q = realloc(dst->field, sizeof(T_) * (size_t)(total ? total : 1));   \
    if (!q) return ORC_ENOMEM;                                                 \
    dst->field = (T_ *)q;                                                      \
    memcpy(dst->field + base, src->field, sizeof(T_) * (size_t)n);             \
  } while (0)
  APPEND2(feature, int32_t);
  APPEND2(threshold, float);
  APPEND2(left, uint32_t);
  APPEND2(right, uint32_t);
  APPEND2(value, uint32_t);
  APPEND2(n_samples, uint32_t);
  APPEND2(depth, uint32_t);
#undef APPEND2
  for (int32_t t = 0; t < src->n_trees; ++t) dst->tree_off[dst->n_trees + t + 1] = base + src->tree_off[t + 1];
  dst->n_trees += src->n_trees;
  return ORC_OK;
}

typedef struct {
  const float *X_T;
  const uint32_t *y;
  int32_t N, F, K;
  const orc_params *prm;
  const uint32_t *sorted;
  int32_t T;
  const uint32_t *boot_ids;
  const int64_t *boot_off;
  orc_forest *parts; /* one forest per tree */
  int32_t next;      /* next tree to grow */
  pthread_mutex_t mu;
} fit_job;

static void *fit_worker(void *arg) {
  fit_job *j = (fit_job *)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    const int32_t t = j->next < j->T ? j->next++ : -1;
    pthread_mutex_unlock(&j->mu);
    if (t < 0) return NULL;
    orc_forest *part = &j->parts[t];
    part->tree_off = (int64_t *)calloc(2, sizeof(int64_t));
    if (!part->tree_off) return (void *)(intptr_t)ORC_ENOMEM;
    int rc = fit_one_tree(j->X_T, j->y, j->N, j->F, j->K, j->prm, j->sorted, t, j->boot_ids, j->boot_off, part);
    if (rc != ORC_OK) return (void *)(intptr_t)rc;
  }
}

int orc_fit(const float *X_T, const uint32_t *y, int32_t N, int32_t F, int32_t K, const uint32_t *sorted,
            int32_t T, const uint32_t *boot_ids, const int64_t *boot_off, const orc_params *prm,
            orc_forest *out) {
  return orc_fit_threads(X_T, y, N, F, K, sorted, T, boot_ids, boot_off, prm, 1, out);
}

int orc_fit_threads(const float *X_T, const uint32_t *y, int32_t N, int32_t F, int32_t K, const uint32_t *sorted,
                    int32_t T, const uint32_t *boot_ids, const int64_t *boot_off, const orc_params *prm,
                    int32_t threads, orc_forest *out) {
  memset(out, 0, sizeof(*out));
  if (N <= 0 || N > ORC_MAX_SAMPLES || F <= 0 || K <= 0 || T <= 0 || !prm) return ORC_EINVAL;
  if (prm->max_features <= 0 || prm->max_features > F) return ORC_EINVAL;
  for (int32_t i = 0; i < N; ++i)
    if (y[i] >= (uint32_t)K) return ORC_EINVAL;
  uint32_t *own = NULL;
  if (!sorted) {
    own = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)F * N);
    if (!own) return ORC_ENOMEM;
    int r = orc_presort(X_T, N, F, own);
    if (r) {
      free(own);
      return r;
    }
    sorted = own;
  }
  if (boot_ids) {
    for (int32_t t = 0; t < T; ++t) {
      int64_t a = boot_off[t], b = boot_off[t + 1];
      if (b <= a) {
        free(own);
        return ORC_EINVAL;
      }
      for (int64_t j = a; j < b; ++j)
        if (boot_ids[j] >= (uint32_t)N || (j > a && boot_ids[j] <= boot_ids[j - 1])) {
          free(own);
          return ORC_EINVAL;
        }
    }
  }
  out->tree_off = (int64_t *)calloc((size_t)T + 1, sizeof(int64_t));
  if (!out->tree_off) {
    free(own);
    return ORC_ENOMEM;
  }
  int rc = ORC_OK;
  if (threads <= 1 || T == 1) {
    for (int32_t t = 0; t < T && rc == ORC_OK; ++t) rc = fit_one_tree(X_T, y, N, F, K, prm, sorted, t, boot_ids,
                                                                   boot_off, out);
  } else {
    /* trees are independent: each worker grows its share into forests of
     * its own, appended in tree order afterwards */
    fit_job job = {X_T, y, N, F, K, prm, sorted, T, boot_ids, boot_off, NULL, 0, PTHREAD_MUTEX_INITIALIZER};
    job.parts = (orc_forest *)calloc((size_t)T, sizeof(orc_forest));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    int32_t started = 0;
    if (!job.parts || !th) rc = ORC_ENOMEM;
    for (int32_t i = 0; i < threads && rc == ORC_OK; ++i) {
      if (pthread_create(&th[i], NULL, fit_worker, &job) != 0) break;
      ++started;
    }
    if (rc == ORC_OK && started == 0) fit_worker(&job); /* no thread to spare: this one */
    for (int32_t i = 0; i < started; ++i) {
      void *r = NULL;
      pthread_join(th[i], &r);
      if (r && rc == ORC_OK) rc = (int)(intptr_t)r;
    }
    for (int32_t t = 0; t < T && rc == ORC_OK && job.parts; ++t)
      if (!job.parts[t].tree_off) rc = ORC_ENOMEM;
    for (int32_t t = 0; t < T && rc == ORC_OK; ++t) rc = append_forest(out, &job.parts[t]);
    if (job.parts)
      for (int32_t t = 0; t < T; ++t) orc_forest_free(&job.parts[t]);
    free(job.parts);
    free(th);
  }
  free(own);
  if (rc != ORC_OK) orc_forest_free(out);
  return rc;
}

void orc_forest_free(orc_forest *f) {
  if (!f) return;
  free(f->tree_off);
  free(f->feature);
  free(f->threshold);
  free(f->left);
  free(f->right);
  free(f->value);
  free(f->n_samples);
  free(f->depth);
  memset(f, 0, sizeof(*f));
}

/* Traversal of predict.cu:50-67 (leaf when a child index is 0, left iff
 * x < threshold) and the vote of random_forest.py:358 (np.argmax of
 * np.bincount: lowest label on ties).  Rows are cast to float (D12). */
int orc_predict(const orc_forest *f, const float *X, int64_t R, int32_t F, int32_t K, uint32_t *out_labels,
                uint32_t *votes) {
  if (!f || K <= 0 || F <= 0) return ORC_EINVAL;
  uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)K);
  if (!cnt) return ORC_ENOMEM;
  for (int64_t r = 0; r < R; ++r) {
    const float *x = X + (size_t)r * F;
    for (int32_t k = 0; k < K; ++k) cnt[k] = 0;
    for (int32_t t = 0; t < f->n_trees; ++t) {
      int64_t base = f->tree_off[t];
      int64_t node = 0;
      while (f->left[base + node] != 0 && f->right[base + node] != 0)
        node = (x[f->feature[base + node]] < f->threshold[base + node]) ? f->left[base + node]
                                                                          : f->right[base + node];
      uint32_t lab = f->value[base + node];
      if (lab < (uint32_t)K) cnt[lab]++;
    }
    out_labels[r] = majority(cnt, (uint32_t)K);
    if (votes)
      for (int32_t k = 0; k < K; ++k) votes[(size_t)r * K + k] = cnt[k];
  }
  free(cnt);
  return ORC_OK;
}

// internal.h -- structures shared by the host driver (api.hip) and the
// gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cta {

constexpr int kBlock = 256;          // threads per block; every kernel works wave by wave
constexpr int kWavesPerBlock = kBlock / 64;
constexpr uint32_t kLeaf = 0xFFFFFFFFu;

// A node waiting for its split (one entry of a level's frontier).
struct FNode {
  uint32_t tree;   // tree index within the resident batch
  uint32_t start;  // segment start in every sorted row of the tree
  uint32_t n;      // samples in the node
  uint32_t nid;    // node id in the batch's node arena
  uint32_t depth;
  uint32_t src;    // rows holding the segment: 0 presorted (shared), 1 buffer 0, 2 buffer 1
  uint32_t pad0, pad1;
};

// Per-node scratch of one level.
struct NAux {
  uint32_t scan_base, scan_ntask, nch_s, fpt;  // split-scan tasks
  uint32_t hist_base;                          // class histograms of multi-chunk scans
  int32_t best_f;                              // -1: no candidate (leaf)
  uint32_t best_p, nL;
  uint32_t left_id, right_id;
  uint32_t flags;                              // kSplit | kLiveL | kLiveR
  uint32_t dst;                                // rows receiving the children
  uint32_t fill_base, nch_f;
  uint32_t resh_big;                           // 1: rows reshuffled by whole workgroups
  uint32_t resh_off;                           // first task in the tree's big or small list
  uint32_t resh_ntask, rpt;                    // small tasks: count and rows per task
  uint32_t pad[2];
};
constexpr uint32_t kSplit = 1u, kLiveL = 2u, kLiveR = 4u;

// A unit of wave work: node, first row/feature slot and count (low/high
// 16 bits), chunk index and the node's chunk count.
struct Task {
  uint32_t node, first_count, chunk, nch;
};

struct ScanRes {
  uint64_t num, den;
  uint32_t pos;
  int32_t feat;
  uint32_t ok, pad;
};

// Counters of one level, in device memory.
struct Ctl {
  uint32_t q, q_next;
  uint32_t n_scan, n_fill, n_hist, n_plan;
  uint32_t n_nodes;
  uint32_t err;
  uint32_t pad[8];
  unsigned long long bytes_scan, bytes_fill, bytes_resh, bytes_boot;
};
constexpr uint32_t kErrCapacity = 1u, kErrBadNode = 2u;

// Flattened forest node for prediction: pre-order, so a split node's left
// child is the next node; leaves have feat = -1 and carry their label.
struct PNode {
  float thr;
  int32_t feat;
  uint32_t right;  // global index of the right child
  uint32_t value;
};

// Node storage of a resident batch of trees, indexed by arena id.
struct Arena {
  int32_t *feat;   // -1 for a leaf
  float *thr;
  uint32_t *left, *right, *value, *nsamp, *depth;
};

template <class IdxT, class LabT>
struct FitArgs {
  const float *X_T;   // F x N, feature-major
  const LabT *y;      // N compact labels
  const IdxT *si0;    // F x N presorted rows
  IdxT *buf0, *buf1;  // per tree F x N ping-pong rows
  uint32_t *bits;     // per tree W words: bit s set = sample s goes left
  uint32_t W;         // words per tree bitmap, (N + 31) / 32
  uint32_t T;         // trees in the batch
  int32_t N, F, K, m, mss, max_depth, identity;
  uint32_t tree_base;  // global index of batch tree 0 (keys the permutation)
  uint64_t seed;
  uint32_t Cs, Cf, Cr;  // chunk sizes: scan, fill, reshuffle
  // level state
  const FNode *front;
  const uint32_t *cnt;  // q x K class counts of the frontier
  FNode *next_front;
  uint32_t *next_cnt;
  uint32_t *cntL;  // q x K left-child counts
  NAux *aux;
  int32_t *feats;  // q x m evaluated features
  Task *scan_tasks;
  ScanRes *res;
  uint32_t *hist;  // class histograms / carries of multi-chunk scans
  Task *fill_tasks;
  // reshuffle plan: per-tree task lists and workgroup ranges
  Task *big_tasks, *small_tasks;
  uint32_t *cnt_big, *cnt_small;  // [T] per-level counters
  unsigned long long *elems;      // [T] elements to move per tree
  uint32_t *big_off, *small_off;  // [T + 1]
  uint32_t *wg_off;               // [T + 1] workgroups per tree (prefix)
  Arena arena;
  Ctl *ctl;
  uint32_t cap_q, cap_scan, cap_hist, cap_fill, cap_big, cap_small, cap_nodes;
};

}  // namespace cta

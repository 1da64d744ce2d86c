// internal.h -- structures shared by the host driver (api.hip) and the
// gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cta {

constexpr int kBlock = 256;          // threads per block; every kernel works wave by wave
constexpr int kWavesPerBlock = kBlock / 64;
constexpr uint32_t kLeaf = 0xFFFFFFFFu;
// Label of a leaf child still to be looked up: its first sample, once the
// side bitmaps are built.
constexpr uint32_t kFirstSample = 0xFFFFFFFEu;

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
  uint32_t flags;                              // kSplit | kLiveL | kLiveR, kCountRl, kInvert
  uint32_t dst;                                // rows receiving the children
  uint32_t fill_base, nch_f;                   // bitmap-fill task slots: first, count (ceil(n / Cf))
  uint32_t resh_big;                           // 1: rows reshuffled by whole workgroups
  uint32_t resh_ntask, rpt;                    // reshuffle tasks of the node, rows per small task
  uint32_t lab[2];                             // labels of leaf children (kLeaf: live child)
  uint32_t nf;                                 // evaluated features that have rows (scan tasks)
  uint32_t rl_pos;                             // best split of the evaluated row-free features:
  uint64_t rl_num, rl_den;                     //   position (zeros - 1), Q numerator / denominator,
  int32_t rl_feat;                             //   feature (-1: none)
  uint32_t pad;
};
constexpr uint32_t kSplit = 1u, kLiveL = 2u, kLiveR = 4u;
// Set by k_decide on a split node with a live child that stays in the levels
// when there are row-free features: its fill tasks count the smaller side.
constexpr uint32_t kCountRl = 8u;
// Set by k_decide on a split node whose right child is the smaller: the side
// bitmaps mark a split node's smaller child, so its marks are the right one's.
constexpr uint32_t kInvert = 16u;

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

// Levels whose frontier size and id count are kept in Ctl.
constexpr uint32_t kLevelRing = 64;
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b % 8), each XCD with its own L2.  Split-scan tasks are kept in one list
// per class, a feature's tasks in one or two classes, and workgroup b works
// on class b % 8, so the value gathers of a column hit few L2s.
constexpr uint32_t kXcd = 8;

// Counters of one level, in device memory.  Nothing is allocated with
// atomics: the totals are written by the offset scans over the frontier.
struct Ctl {
  uint32_t q;        // frontier size
  uint32_t n_nodes;  // arena ids in use
  uint32_t err;
  uint32_t depth;      // levels completed (ends of a level: ctl_next_body)
  uint32_t n_plan[2];  // reshuffle plan entries: big tasks, small tasks
  uint32_t pad1[2];
  uint32_t n_scan, n_hist, n_fill, n_cls[kXcd], pad2;    // selection-scan totals
  uint32_t c_ids, c_live, c_big, c_small, c_tiles, c_sub, pad3[2];  // children-scan totals
  unsigned long long bytes_scan, bytes_fill, bytes_resh, bytes_boot, bytes_resh_small, bytes_rl, bytes_sub,
      bytes_scan_skip;
  // small subtrees: nodes queued so far (whole levels), the next one a
  // subtree wave takes, and the deepest node depth the subtree waves made
  uint32_t n_sub, sub_next, sub_depth, pad4;
  // after level `depth`: frontier size and ids in use, at depth % kLevelRing
  uint32_t lv_q[kLevelRing], lv_nodes[kLevelRing];
};
// Fields of the two per-node offset scans.
constexpr int kSelFields = 3 + kXcd;  // scan tasks, histogram words, fill slots, scan tasks per class
// child ids, next-frontier slots, big / small reshuffle tasks, big tiles,
// small-subtree queue slots
constexpr int kChFields = 6;
// Algorithmic-byte counters: per-wave sums flushed into kByteSlots spread
// slots (one 128-byte line each) and folded into Ctl at the end of a level.
constexpr uint32_t kByteSlots = 64, kByteStride = 16;
enum ByteKind {
  kBytesScan = 0,
  kBytesFill = 1,
  kBytesResh = 2,
  kBytesBoot = 3,
  kBytesReshSmall = 4,
  kBytesRl = 5,
  kBytesSub = 6,
  kBytesScanSkip = 7  // split-scan bytes of chunks skipped by the score bound
};
constexpr uint32_t kByteKinds = 8;
constexpr uint32_t kErrCapacity = 1u, kErrBadNode = 2u;

// Flattened forest node for prediction: pre-order, so a split node's left
// child is the next node; leaves have feat = -1 and carry their label.
struct PNode {
  float thr;
  int32_t feat;
  uint32_t right;  // global index of the right child
  uint32_t value;
};

// Predict table in blocks: a node at depth 3k of its tree roots a block that
// holds its subtree down to depth 3k + 2 in heap order (slot 0 the root,
// slots 2s + 1 and 2s + 2 the children of slot s; slot 7 unused).  A block is
// one 128-byte line, so a walk fetches one line per three levels and takes
// the children inside a block from that line.
struct BNode {
  float thr;
  int32_t feat;   // -1: leaf
  uint32_t a, b;  // leaf: a = label; split node in slots 3..6: blocks of its left / right child
};
constexpr uint32_t kBlockSlots = 8;

// Node storage of a resident batch of trees, indexed by arena id.
struct Arena {
  int32_t *feat;   // -1 for a leaf
  float *thr;
  uint32_t *left, *right, *value, *nsamp, *depth;
};

// Row entries: 16-bit sample ids (N <= 65,535), or 32-bit words holding
// the sample id in the low kIdBits bits and its label in the high bits
// (N < 2^22, K <= 1,024), so the split scan reads labels with the ids.
constexpr uint32_t kIdBits = 22, kIdMask = (1u << kIdBits) - 1u;

// fcode[f]: byte offset of feature f's code column (a multiple of 256) and
// the column kind (0: none, read X_T; 1: bits, 2: u8, 3: u16).
__host__ __device__ __forceinline__ uint32_t fc_kind(uint64_t fc) { return (uint32_t)(fc & 3u); }
__host__ __device__ __forceinline__ uint64_t fc_off(uint64_t fc) { return fc & ~(uint64_t)255; }

template <class IdxT>
__host__ __device__ __forceinline__ uint32_t sample_of(IdxT v) {
  if constexpr (sizeof(IdxT) == 4) return (uint32_t)v & kIdMask;
  else return (uint32_t)v;
}

template <class IdxT, class LabT>
struct FitArgs {
  const float *X_T;   // F x N, feature-major
  const LabT *y;      // N compact labels
  const IdxT *si0;    // F x N presorted rows
  const uint8_t *codes;    // value code columns of the features that have one
  const uint64_t *fcode;   // [F] code column byte offset | kind
  // Row-free features: two-valued features other than feature 0 keep no
  // rows.  Their one candidate split, their constancy and their side bits
  // follow from per-node, per-class counts of the samples holding the
  // higher value (rlc), which a pass over each split node's samples
  // (through its feature-0 row and the per-sample bit signatures) produces
  // for the children.
  uint32_t Frl, Wsig;       // row-free features, 32-bit signature words per sample
  uint32_t Fr;              // features with rows
  const int32_t *rl_of;     // [F] row-free index of each feature, or -1
  const int32_t *slot_of;   // [F] row slot of each feature, or -1
  const int32_t *row_feat;  // [Fr] feature of each row slot
  const float *rl_thr;      // [Frl] threshold between the two values
  const uint32_t *sig;      // [N x Wsig] bit j: sample holds the higher value of row-free feature j
  const uint32_t *rlc;      // frontier: [q x Frl x K] samples of class k holding the higher value
  uint32_t *rlc_next;       // the same for the next frontier
  uint32_t *rl_small;       // frontier: [q x Frl x K] counts of each split node's smaller side
  IdxT *buf0, *buf1;  // per tree Fr x N ping-pong rows (row slots)
  uint32_t *bits;     // per tree W words: bit s set = sample s goes to its node's smaller child
                      // (left, or right under kInvert; in-bag marks for the bootstrap compaction)
  uint32_t W;         // words per tree bitmap, (N + 31) / 32
  uint32_t T;         // trees in the batch
  int32_t N, F, K, m, mss, max_depth, identity;
  uint32_t tree_base;  // global index of batch tree 0 (keys the permutation)
  uint64_t seed;
  uint32_t Cs, Cf, Cr;  // chunk sizes: scan, fill, reshuffle
  uint32_t hist_lds;    // 1: split-scan class counts in per-wave LDS counters even when K <= 64
  // Live children of at most sub_max samples (0: none) leave the levels for
  // the small-subtree queue; a subtree wave grows each one to its leaves.
  uint32_t sub_max;
  FNode *sub_queue;
  uint32_t cap_sub;
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
  // per-node offset scans: counts and exclusive offsets (kSelFields or
  // kChFields per node)
  uint32_t *vals, *offs;
  uint32_t *tlo, *thi;  // [T] frontier range of each tree (frontiers are sorted by tree)
  // reshuffle plan: per-tree task ranges and workgroup ranges
  Task *big_tasks, *small_tasks;
  uint32_t *plan;    // [T x 4] big begin/end, small begin/end
  uint32_t *wg_off;  // [2 x (T + 1)] workgroups per tree (prefix): big tasks, then small tasks
  unsigned long long *bslot;  // [kByteKinds x kByteSlots x kByteStride] byte counters
  Arena arena;
  Ctl *ctl;
  Ctl *ctl_host;  // when set, the end of the level also copies Ctl here (host-mapped memory)
  uint32_t cap_q, cap_scan, cap_hist, cap_fill, cap_big, cap_small, cap_nodes;

  // label of a row entry: carried in the word, or looked up by sample id
  __device__ __forceinline__ uint32_t label(IdxT v) const {
    if constexpr (sizeof(IdxT) == 4) return (uint32_t)v >> kIdBits;
    else return (uint32_t)y[v];
  }
};

}  // namespace cta

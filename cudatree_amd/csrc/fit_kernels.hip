// fit_kernels.hip -- level-synchronous random-forest growth on gfx950.
//
// One level of every resident tree is processed by a fixed sequence of
// launches.  The per-node kernels are wave-granular: a 64-lane wave takes one
// frontier node or one task (a chunk of a node's segment in one or more
// sorted rows) at a time from a grid-stride loop whose bound is read from
// device memory, so the host never has to size grids from device-side
// counts.  Lists are laid out by exclusive scans over per-node counts, never
// by atomics, so frontiers stay sorted by tree and every tree's nodes and
// tasks are contiguous.
//
//   k_select      per node: feature visiting order, constant-feature skip,
//                 task / histogram / fill-slot counts  (bfs_module.cu:122-139)
//   k_off_*       exclusive offsets of per-node counts
//   k_select_emit per node: split-scan task list
//   k_scan_hist   per multi-chunk task: class histogram of the chunk
//                                                (dfs_module.cu:24-80)
//   k_scan_carry  per multi-chunk feature: exclusive prefix over chunks
//                                                (dfs_module.cu:82-106)
//   k_scan        per task: Gini scan of the chunk, best position
//                                                (dfs_module.cu:310-443, bfs_module.cu:65-203)
//   k_decide      per node: best split over tasks, threshold, left-child
//                 class counts, fill tasks      (dfs_module.cu:446-491, 158-208, bfs_module.cu:206-235, 379-401),
//                 leaf tests of both children (random_tree.py:71-165), child
//                 and reshuffle counts
//   k_fill_lds    per tree: side bitmap built in LDS
//                                                (dfs_module.cu:210-221, bfs_module.cu:238-264)
//   k_children_emit per split node: ids, next frontier, reshuffle tasks
//   k_plan / k_resh
//                 stable partition of every row (dfs_module.cu:224-307, bfs_module.cu:266-377)
//
// The split scan parallelises the reference's serial recurrence
// S_L(p) = S_L(p-1) + 2 c_{y_p}(p-1) + 1 (dfs_module.cu:396-397): each lane's
// c_{y_p}(p-1) is the chunk carry of its label plus its rank among earlier
// lanes with the same label (one ballot per distinct label in the wave), and
// S_L and D = sum_k tot_k c_k are then wave prefix sums.
#include "device_util.h"
#include "internal.h"

namespace cta {

#define WAVE_LOOP(count_expr)                                                        \
  const uint32_t _wid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;           \
  const uint32_t _nw = gridDim.x * kWavesPerBlock;                                   \
  const uint32_t _count = (count_expr);                                              \
  for (uint32_t item = _wid; item < _count; item += _nw)

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// Rows holding a node's segment: src 0 the presorted rows (one per feature,
// shared by the trees), 1 / 2 the tree's row slots in buffer 0 / 1.
template <class IdxT, class LabT>
__device__ __forceinline__ const IdxT *rows_of_slot(const FitArgs<IdxT, LabT> &a, uint32_t src, uint32_t tree,
                                                    uint32_t s) {
  if (src == 0) return a.si0 + (size_t)a.row_feat[s] * a.N;
  const IdxT *b = (src == 1) ? a.buf0 : a.buf1;
  return b + ((size_t)tree * a.Fr + s) * a.N;
}
// the row of feature f (a feature with rows)
template <class IdxT, class LabT>
__device__ __forceinline__ const IdxT *rows_of(const FitArgs<IdxT, LabT> &a, uint32_t src, uint32_t tree,
                                               uint32_t f) {
  if (src == 0) return a.si0 + (size_t)f * a.N;
  const IdxT *b = (src == 1) ? a.buf0 : a.buf1;
  return b + ((size_t)tree * a.Fr + (uint32_t)a.slot_of[f]) * a.N;
}

// Split-scan tasks aimed for at a level whose frontier is small.
#ifndef CTA_SCAN_TASK_TARGET
#define CTA_SCAN_TASK_TARGET 4096
#endif
constexpr uint32_t kScanTaskTarget = CTA_SCAN_TASK_TARGET;

// Split-scan class (< kXcd) of the tasks of feature f in frontier node i.
// Workgroups are dispatched to the XCDs round-robin (workgroup b on XCD
// b % kXcd, each XCD with its own L2) and k_scan's workgroup b takes class
// b % kXcd.  Row slot s owns kXcd units, s kXcd .. s kXcd + kXcd - 1, the
// node picks one of them, and class x holds units x Fr .. x Fr + Fr - 1: the
// classes get equal shares of the work and a feature's tasks span at most
// 1 + ceil(kXcd / Fr) classes, so its gathered column sits in few L2s.
template <class IdxT, class LabT>
__device__ __forceinline__ uint32_t scan_cls(const FitArgs<IdxT, LabT> &a, uint32_t f, uint32_t i) {
  return ((uint32_t)a.slot_of[f] * kXcd + i % kXcd) / a.Fr;
}

template <class IdxT, class LabT>
__device__ __forceinline__ IdxT *rows_out(const FitArgs<IdxT, LabT> &a, uint32_t dst, uint32_t tree,
                                          uint32_t s) {
  IdxT *b = (dst == 1) ? a.buf0 : a.buf1;
  return b + ((size_t)tree * a.Fr + s) * a.N;
}

__device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint32_t smem_wave_offset(int K) { return (threadIdx.x / kWave) * 2u * (uint32_t)K; }

// Adds a wave's algorithmic byte count to the counter of `kind` (one
// atomic per wave and launch, spread over kByteSlots lines).
template <class IdxT, class LabT>
__device__ __forceinline__ void flush_bytes(const FitArgs<IdxT, LabT> &a, ByteKind kind, unsigned long long v) {
  if (lane_id() == 0 && v)
    atomicAdd(&a.bslot[((uint32_t)kind * kByteSlots + blockIdx.x % kByteSlots) * kByteStride], v);
}

// Distinct-value index of sample s in a value code column of kind 1 (bits),
// 2 (bytes) or 3 (16-bit).
template <uint32_t kKind>
__device__ __forceinline__ uint32_t code_at_k(const uint8_t *cb, uint32_t s) {
  if constexpr (kKind == 1) return (reinterpret_cast<const uint32_t *>(cb)[s >> 5] >> (s & 31u)) & 1u;
  else if constexpr (kKind == 2) return cb[s];
  else return reinterpret_cast<const uint16_t *>(cb)[s];
}
__device__ __forceinline__ uint32_t code_at(const uint8_t *cb, uint32_t kind, uint32_t s) {
  if (kind == 1) return code_at_k<1>(cb, s);
  if (kind == 2) return code_at_k<2>(cb, s);
  return code_at_k<3>(cb, s);
}

// Row entries at positions pb + u * kWave + lane (u < U) of a row whose valid
// positions end at p1 > pb, loaded unconditionally (positions past p1 - 1
// load the last entry; callers mask them), so that the U loads and what
// depends on them issue together rather than one masked load at a time.
template <int U, class IdxT>
__device__ __forceinline__ void load_row_batch(const IdxT *row, uint32_t pb, uint32_t p1, uint32_t (&e)[U]) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int u = 0; u < U; ++u) e[u] = (uint32_t)row[min(pb + (uint32_t)u * kWave + lane, p1 - 1u)];
}

// ------------------------------------------------------------------------
// Offset scans.  Per-node counts (NF fields, node-major) become exclusive
// offsets and totals, so every list a level builds (split-scan tasks,
// histograms, fill tasks, child ids, next frontier, reshuffle tasks) is laid
// out in frontier order without atomics.  Frontiers are sorted by tree and
// children are emitted in parent order, so each tree's nodes and tasks stay
// contiguous.
// ------------------------------------------------------------------------
constexpr int kOffThreads = 256, kOffItems = 4;
constexpr uint32_t kOffTile = kOffThreads * kOffItems;  // items per block of the multi-block scan
constexpr int kOffSingleThreads = 1024;

// Per-block sums of each field.
template <int NF>
__global__ __launch_bounds__(kOffThreads) void k_off_sums(const uint32_t *v, const uint32_t *qp, uint32_t *bsum) {
  __shared__ uint32_t ws[kOffThreads / kWave][NF];
  const uint32_t q = *qp;
  if (blockIdx.x * kOffTile >= q) return;  // grid sized for a bound on q
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const uint32_t i0 = blockIdx.x * kOffTile + tid * kOffItems;
  uint32_t s[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) s[f] = 0;
#pragma unroll
  for (int k = 0; k < kOffItems; ++k)
    if (i0 + k < q)
#pragma unroll
      for (int f = 0; f < NF; ++f) s[f] += v[(size_t)(i0 + k) * NF + f];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s[f] = wave_sum32(s[f]);
    if (lane == 0) ws[wave][f] = s[f];
  }
  __syncthreads();
  if (tid < (uint32_t)NF) {
    uint32_t t = 0;
    for (int w = 0; w < kOffThreads / kWave; ++w) t += ws[w][tid];
    bsum[(size_t)blockIdx.x * NF + tid] = t;
  }
}

// One block: exclusive scan of n items (in place allowed), totals to tot.
template <int NF>
__global__ __launch_bounds__(kOffSingleThreads) void k_off_single(const uint32_t *v, const uint32_t *np,
                                                                 uint32_t ndiv, uint32_t *out, uint32_t *tot) {
  constexpr int kW = kOffSingleThreads / kWave;
  const uint32_t n = (*np + ndiv - 1) / ndiv;  // items, or blocks of ndiv items
  __shared__ uint32_t ws[kW][NF];
  __shared__ uint32_t carry[NF];
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  if (tid < (uint32_t)NF) carry[tid] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += kOffSingleThreads * kOffItems) {
    const uint32_t i0 = base + tid * kOffItems;
    uint32_t x[kOffItems][NF], ts[NF], inc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) ts[f] = 0;
#pragma unroll
    for (int k = 0; k < kOffItems; ++k)
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        x[k][f] = (i0 + k < n) ? v[(size_t)(i0 + k) * NF + f] : 0u;
        ts[f] += x[k][f];
      }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      inc[f] = wave_incl_scan(ts[f]);
      if (lane == kWave - 1) ws[wave][f] = inc[f];
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      // the waves before this one: an exclusive scan of the kW wave sums
      const uint32_t wsum = lane < (uint32_t)kW ? ws[lane][f] : 0u;
      const uint32_t wincl = wave_incl_scan(wsum);
      uint32_t run = carry[f] + readlane(wincl - wsum, wave) + inc[f] - ts[f];
#pragma unroll
      for (int k = 0; k < kOffItems; ++k) {
        if (i0 + k < n) out[(size_t)(i0 + k) * NF + f] = run;
        run += x[k][f];
      }
    }
    __syncthreads();
    if (tid < (uint32_t)NF) {
      uint32_t t = 0;
      for (int w = 0; w < kW; ++w) t += ws[w][tid];
      carry[tid] += t;
    }
    __syncthreads();
  }
  if (tid < (uint32_t)NF) tot[tid] = carry[tid];
}

// Exclusive offsets of one block's items given the scanned block sums.
template <int NF>
__global__ __launch_bounds__(kOffThreads) void k_off_apply(const uint32_t *v, const uint32_t *qp,
                                                           const uint32_t *bsum, uint32_t *out) {
  __shared__ uint32_t ws[kOffThreads / kWave][NF];
  const uint32_t q = *qp;
  if (blockIdx.x * kOffTile >= q) return;  // grid sized for a bound on q
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const uint32_t i0 = blockIdx.x * kOffTile + tid * kOffItems;
  uint32_t x[kOffItems][NF], ts[NF], inc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) ts[f] = 0;
#pragma unroll
  for (int k = 0; k < kOffItems; ++k)
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      x[k][f] = (i0 + k < q) ? v[(size_t)(i0 + k) * NF + f] : 0u;
      ts[f] += x[k][f];
    }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    inc[f] = wave_incl_scan(ts[f]);
    if (lane == kWave - 1) ws[wave][f] = inc[f];
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    uint32_t run = bsum[(size_t)blockIdx.x * NF + f] + inc[f] - ts[f];
    for (uint32_t w = 0; w < wave; ++w) run += ws[w][f];
#pragma unroll
    for (int k = 0; k < kOffItems; ++k) {
      if (i0 + k < q) out[(size_t)(i0 + k) * NF + f] = run;
      run += x[k][f];
    }
  }
}

// ------------------------------------------------------------------------
// k_select: visiting order of every frontier node, its task and slot counts
// (scan tasks, histogram words, fill-task slots) and the frontier range of
// each tree.
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_select(FitArgs<IdxT, LabT> a) {
  const uint32_t lane = lane_id();
  const uint32_t m = (uint32_t)a.m, F = (uint32_t)a.F;
  const uint32_t q = a.ctl->q;
  // the side bitmaps start the level clear (the fill ORs into them; the
  // previous level's readers are done)
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < (size_t)a.T * a.W;
       w += (size_t)gridDim.x * blockDim.x)
    a.bits[w] = 0u;
  unsigned long long bytes = 0;
  WAVE_LOOP(q) {
    const uint32_t i = item;
    const FNode nd = a.front[i];
    const uint32_t n = nd.n;
    const uint32_t nch = ceil_div(n, a.Cs);
    Perm pm = make_perm(a.seed, a.tree_base + nd.tree, nd.depth, nd.start, F, a.identity != 0);
    // tasks: one per (feature, chunk) of a multi-chunk node; a single-chunk
    // node's features grouped about Cs positions per task.  A task's class is
    // scan_cls of its (first) feature; lane c counts the tasks of class c.
    // (a small frontier's nodes grouped less, so that the level still has
    // about kScanTaskTarget tasks to spread over the waves)
    const uint32_t fpt =
        (nch > 1) ? 1u : max(1u, (uint32_t)min<uint64_t>(a.Cs / n, (uint64_t)q * m / kScanTaskTarget));
    const uint32_t wt = (nch > 1) ? nch : 1u;
    const uint32_t K = (uint32_t)a.K;
    uint32_t ncls = 0;
    uint32_t nf = 0;    // evaluated features with rows (they get scan tasks)
    uint32_t nsel = 0;  // evaluated features
    Cand rlb = {0, 1, 0, -1, 0};  // best candidate of the evaluated row-free features
    const uint32_t *tc = a.cnt + (size_t)i * K;  // the node's class counts
    for (uint32_t k0 = 0; k0 < F && nsel < m; k0 += kWave) {
      const uint32_t k = k0 + lane;
      bool nonconst = false;
      uint32_t f = 0;
      int32_t jr = -1;
      Cand cd = {0, 1, 0, -1, 0};  // a row-free feature's one candidate
      if (k < F) {
        f = perm_at(pm, k);
        // the feature's tables read together
        jr = a.Frl ? a.rl_of[f] : -1;
        const int32_t slot = a.slot_of[f];
        const uint64_t fc = a.fcode[f];
        if (jr >= 0) {
          // row-free: constant when no or every sample holds the higher
          // value; its one candidate sends the samples holding the lower
          // value left (zeros = n - ones), scored with the integers of the
          // full scan.  The class counts are loaded eight at a time.
          const uint32_t *oc = a.rlc + ((size_t)i * a.Frl + (uint32_t)jr) * K;
          uint64_t SL = 0, Dd = 0, T2 = 0;
          uint32_t ones = 0;
          for (uint32_t c0 = 0; c0 < K; c0 += 8) {
            uint32_t o[8], t[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
              o[u] = c0 + u < K ? oc[c0 + u] : 0u;
              t[u] = c0 + u < K ? tc[c0 + u] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
              const uint64_t cl = t[u] - o[u];
              SL += cl * cl;
              Dd += (uint64_t)t[u] * cl;
              T2 += (uint64_t)t[u] * t[u];
              ones += o[u];
            }
          }
          nonconst = ones != 0u && ones != n;
          const uint64_t nL = n - ones, nR = ones;
          const uint64_t SR = T2 + SL - 2 * Dd;
          cd = {SL * nR + SR * nL, nL * nR, (uint32_t)nL - 1u, (int32_t)f, 1u};
        } else {
          const IdxT *row = (nd.src == 0 ? a.si0 + (size_t)f * a.N
                                         : (nd.src == 1 ? a.buf0 : a.buf1) +
                                               ((size_t)nd.tree * a.Fr + (uint32_t)slot) * a.N) +
                            nd.start;
          const uint32_t s0 = sample_of(row[0]), s1 = sample_of(row[nd.n - 1]);
          // constant in the node: first and last sorted values equal (value
          // codes when the feature has them, else the floats)
          const uint32_t kind = fc_kind(fc);
          if (kind == 0) {
            const float *x = a.X_T + (size_t)f * a.N;
            nonconst = x[s0] != x[s1];
          } else {
            nonconst = code_at(a.codes + fc_off(fc), kind, s0) != code_at(a.codes + fc_off(fc), kind, s1);
          }
        }
      }
      const uint64_t mk = ballot(nonconst);
      const uint32_t r = mbcnt(mk);
      const bool take = nonconst && nsel + r < m;
      // features with rows in visiting order; row-free ones scored here
      const bool trow = take && jr < 0;
      const uint64_t mrow = ballot(trow);
      const uint32_t rr = mbcnt(mrow);
      if (trow) a.feats[(size_t)i * m + nf + rr] = (int32_t)f;
      const uint32_t c = trow && (nf + rr) % fpt == 0 ? scan_cls(a, f, i) : kXcd;
#pragma unroll
      for (uint32_t x = 0; x < kXcd; ++x) {
        const uint32_t cnt = (uint32_t)__popcll(ballot(c == x));
        if (lane == x) ncls += cnt * wt;
      }
      if (take && jr >= 0 && cand_better(cd, rlb)) rlb = cd;
      nf += (uint32_t)__popcll(mrow);
      nsel += min((uint32_t)__popcll(mk), m - nsel);
    }
    rlb = wave_best(rlb);
    const uint32_t ntask = (nf == 0) ? 0u : (nch > 1 ? nf * nch : ceil_div(nf, fpt));
    if (lane < kXcd) a.vals[(size_t)i * kSelFields + 3 + lane] = ncls;
    if (lane == 0) {
      uint32_t *v = a.vals + (size_t)i * kSelFields;
      v[0] = ntask;
      v[1] = (nch > 1) ? nf * nch * (uint32_t)a.K : 0u;
      v[2] = ceil_div(n, a.Cf);
      NAux ax = {};
      ax.scan_ntask = ntask;
      ax.nch_s = nch;
      ax.fpt = fpt;
      ax.nf = nf;
      ax.best_f = -1;
      ax.rl_feat = rlb.ok ? rlb.feat : -1;
      ax.rl_pos = rlb.pos;
      ax.rl_num = rlb.num;
      ax.rl_den = rlb.den;
      a.aux[i] = ax;
      if (i == 0 || a.front[i - 1].tree != nd.tree) a.tlo[nd.tree] = i;
      if (i + 1 == q || a.front[i + 1].tree != nd.tree) a.thi[nd.tree] = i + 1;
      // class-totals pass n (b + 1) plus per evaluated feature n (b + 1 + 4):
      // index read, label and value gathers at element size
      bytes += (unsigned long long)n * (sizeof(IdxT) + sizeof(LabT)) +
               (unsigned long long)nf * n * (sizeof(IdxT) + sizeof(LabT) + 4);
    }
  }
  flush_bytes(a, kBytesScan, bytes);
}

// Writes the split-scan tasks of every node at its scanned offset.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_select_emit(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  const uint32_t lane = lane_id();
  // class lists one after another in scan_tasks; lane c < kXcd: start of class c
  uint32_t cls_base = 0;
  {
    const uint32_t cn = lane < kXcd ? a.ctl->n_cls[lane] : 0u;
    cls_base = wave_incl_scan(cn) - cn;
  }
  const uint32_t n_scan = a.ctl->n_scan;
  WAVE_LOOP(a.ctl->q) {
    const uint32_t i = item;
    NAux ax = a.aux[i];
    const uint32_t *o = a.offs + (size_t)i * kSelFields;
    const uint32_t *v = a.vals + (size_t)i * kSelFields;
    const uint32_t base = o[0], hbase = o[1], fbase = o[2];
    const uint32_t ntask = ax.scan_ntask, nch = ax.nch_s, fpt = ax.fpt, nf = ax.nf;
    const bool fits = (uint64_t)n_scan <= a.cap_scan && (uint64_t)hbase + v[1] <= a.cap_hist &&
                      (uint64_t)fbase + v[2] <= a.cap_fill;
    if (!fits) {
      if (lane == 0) {
        atomicOr(&a.ctl->err, kErrCapacity);
        ax.scan_ntask = 0;
        ax.nch_f = 0;
        a.aux[i] = ax;
      }
      continue;
    }
    // lane c < kXcd: where this node's tasks of class c start
    uint32_t cpos = lane < kXcd ? cls_base + o[3 + lane] : 0u;
    // items: features of a multi-chunk node (nch tasks each), else tasks
    const uint32_t nitem = nch > 1 ? nf : ntask;
    for (uint32_t c0 = 0; c0 < nitem; c0 += kWave) {
      const uint32_t it = c0 + lane;
      const bool valid = it < nitem;
      const uint32_t c = valid ? scan_cls(a, (uint32_t)a.feats[(size_t)i * a.m + (nch > 1 ? it : it * fpt)], i) : kXcd;
      uint32_t pos = 0;
#pragma unroll
      for (uint32_t x = 0; x < kXcd; ++x) {
        const uint64_t mk = ballot(c == x);
        const uint32_t b = readlane(cpos, x);
        if (c == x) pos = b + mbcnt(mk) * (nch > 1 ? nch : 1u);
        if (lane == x) cpos += (uint32_t)__popcll(mk) * (nch > 1 ? nch : 1u);
      }
      if (nch == 1) {
        if (valid) {
          Task tk;
          tk.node = i;
          const uint32_t first = it * fpt;
          tk.first_count = first | (min(fpt, nf - first) << 16);
          tk.chunk = 0;
          tk.nch = 1;
          a.scan_tasks[pos] = tk;
        }
      } else {
        // the chunk tasks of this group's features, nch in a row per feature
        const uint32_t cnt = min((uint32_t)kWave, nitem - c0);
        for (uint32_t k0 = 0; k0 < cnt * nch; k0 += kWave) {
          const uint32_t k = k0 + lane;
          const uint32_t jl = min(k / nch, cnt - 1u);
          const uint32_t p = __shfl(pos, (int)jl);
          if (k < cnt * nch) {
            Task tk;
            tk.node = i;
            tk.first_count = (c0 + jl) | (1u << 16);
            tk.chunk = k % nch;
            tk.nch = nch;
            a.scan_tasks[p + tk.chunk] = tk;
          }
        }
      }
    }
    if (lane == 0) {
      ax.scan_base = base;
      ax.hist_base = hbase;
      ax.fill_base = fbase;
      ax.nch_f = v[2];
      a.aux[i] = ax;
    }
  }
}

// ------------------------------------------------------------------------
// k_scan_hist: class histogram of one chunk of one feature row (multi-chunk
// nodes only).  The prefix over chunks gives each chunk its starting counts,
// the decomposition of scan_gini_large + scan_reduce (dfs_module.cu:24-106).
// ------------------------------------------------------------------------
// positions per lane in flight in a chunk-histogram task
#ifndef CTA_HIST_U
#define CTA_HIST_U 8
#endif
constexpr int kHistU = CTA_HIST_U;

template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_scan_hist(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  uint32_t *h = smem + smem_wave_offset(a.K);
  WAVE_LOOP(a.ctl->n_scan) {
    const Task tk = a.scan_tasks[item];
    if (tk.nch <= 1) continue;
    const FNode nd = a.front[tk.node];
    const NAux ax = a.aux[tk.node];
    const uint32_t j = tk.first_count & 0xFFFFu;
    const uint32_t f = (uint32_t)a.feats[(size_t)tk.node * a.m + j];
    const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
    // K <= 64: lane k counts class k in a register, one ballot per distinct
    // label of a tile; else LDS atomics
    const bool regs = K <= (uint32_t)kWave && !a.hist_lds;
    uint32_t hv = 0;
    if (!regs) {
      for (uint32_t k = lane; k < K; k += kWave) h[k] = 0;
      wave_sync();
    }
    const uint32_t p0 = tk.chunk * a.Cs, p1 = min(nd.n, p0 + a.Cs);
    for (uint32_t pb = p0; pb < p1; pb += kHistU * kWave) {
      uint32_t sid[kHistU];
      load_row_batch(row, pb, p1, sid);
      uint32_t l[kHistU];
#pragma unroll
      for (int u = 0; u < kHistU; ++u) l[u] = a.label((IdxT)sid[u]);
#pragma unroll
      for (int u = 0; u < kHistU; ++u)
        if (pb + (uint32_t)u * kWave + lane >= p1) l[u] = 0xFFFFFFFFu;  // past the chunk
      if (regs) {
#pragma unroll
        for (int u = 0; u < kHistU; ++u) {
          uint64_t rem = ballot(l[u] != 0xFFFFFFFFu);
          while (rem) {
            const uint32_t L = readlane(l[u], (uint32_t)__ffsll((unsigned long long)rem) - 1u);
            const uint64_t mm = ballot(l[u] == L);
            if (lane == L) hv += (uint32_t)__popcll(mm);
            rem &= ~mm;
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < kHistU; ++u)
          if (l[u] != 0xFFFFFFFFu) atomicAdd(&h[l[u]], 1u);
      }
    }
    uint32_t *dst = a.hist + ax.hist_base + ((size_t)j * tk.nch + tk.chunk) * K;
    if (regs) {
      if (lane < K) dst[lane] = hv;
    } else {
      wave_sync();
      for (uint32_t k = lane; k < K; k += kWave) dst[k] = h[k];
      wave_sync();
    }
  }
}

// Result-entry mark of a chunk that k_scan_carry has ruled out.
constexpr uint32_t kScanSkipped = 1u;
// Chunks are ruled out for up to kPruneMaxK classes; the chunk boundaries
// of nodes of at least kBoundMinN samples join the known candidates (below
// that, a node's chunks are too few for their value gathers to pay).
constexpr uint32_t kPruneMaxK = 8;
#ifndef CTA_BOUND_MIN_N
#define CTA_BOUND_MIN_N 32768
#endif
constexpr uint32_t kBoundMinN = CTA_BOUND_MIN_N;

// Exclusive prefix over the chunks of each multi-chunk (node, feature);
// each chunk's result entry is cleared on the way, and marked skipped when
// no position of the chunk can hold the node's split.  A split needs a
// strictly higher score than any other candidate of the node, and two are
// known here: the node's best row-free candidate (scored exactly by the
// selection) and, in nodes of at least kBoundMinN samples, the feature's
// chunk boundaries whose two values pass the candidate test (their left
// class counts are the carries).  For p in
// [p0, p1) the left class counts lie between the chunk's carries and the
// next chunk's (sum of squares at most SLe, with nL >= p0 + 1) and the
// right counts only shrink (sum of squares at most SRs = sum_k (t_k -
// c_k)^2 at p0, with nR >= n - p1), so no position of the chunk scores above
// SLe / (p0 + 1) + SRs / (n - p1).  A chunk whose bound is below a known
// candidate's score is skipped; the chunk holding that candidate is not
// (its bound is at least the candidate's score).  The test keeps a relative
// margin of 2^-30 on each side, far above the rounding of these doubles
// (counts and sums of squares below 2^53, scores' numerators below 2^64).
// It is made for up to kC classes, whose counts the prefix pass holds in
// registers; while the prefix pass runs, a chunk's result entry holds its
// bound.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_scan_carry(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  constexpr uint32_t kC = 8;  // classes in flight
  static_assert(kPruneMaxK <= kC, "the bounds take every class from one prefix pass");
  unsigned long long skipped = 0;  // algorithmic split-scan bytes of the chunks ruled out
  WAVE_LOOP(a.ctl->n_scan) {
    const Task tk = a.scan_tasks[item];
    if (tk.nch <= 1 || tk.chunk != 0) continue;
    const NAux ax = a.aux[tk.node];
    const uint32_t j = tk.first_count & 0xFFFFu;
    uint32_t *base = a.hist + ax.hist_base + (size_t)j * tk.nch * K;
    ScanRes *rs = a.res + ax.scan_base + (size_t)j * tk.nch;
    const int32_t f = a.feats[(size_t)tk.node * a.m + j];
    // prune: chunks are tested against the known candidates; bounds: the
    // chunk boundaries join them, and the marks wait for a second pass
    const FNode nd = a.front[tk.node];
    const uint32_t n = nd.n;
    const bool bounds = K <= kPruneMaxK && n >= kBoundMinN;
    const bool prune = K <= kPruneMaxK && (bounds || ax.rl_feat >= 0);
    uint32_t t[kC] = {};
    uint64_t T2 = 0;
    uint32_t kind = 0;
    double known = 0.0;  // the best score of the candidates known so far
    const IdxT *row = nullptr;
    const uint8_t *cb = nullptr;
    const float *x = nullptr;
    if (prune) {
#pragma unroll
      for (uint32_t c = 0; c < kC; ++c) {
        t[c] = c < K ? a.cnt[(size_t)tk.node * K + c] : 0u;
        T2 += (uint64_t)t[c] * t[c];
      }
      if (ax.rl_feat >= 0) known = (double)ax.rl_num / (double)ax.rl_den;
    }
    if (bounds) {
      row = rows_of(a, nd.src, nd.tree, (uint32_t)f) + nd.start;
      const uint64_t fc = a.fcode[f];
      kind = uniform(fc_kind(fc));
      cb = a.codes + fc_off(fc);
      x = a.X_T + (size_t)f * a.N;
    }
    // per class, the chunk histograms become exclusive prefix sums over
    // the chunks: lanes take 64 chunks at a time, kC classes at once
    for (uint32_t k0 = 0; k0 < K; k0 += kC) {
      uint32_t carry[kC];
#pragma unroll
      for (uint32_t c = 0; c < kC; ++c) carry[c] = 0;
      for (uint32_t c0 = 0; c0 < tk.nch; c0 += kWave) {
        const uint32_t ch = c0 + lane;
        uint32_t v[kC];
#pragma unroll
        for (uint32_t c = 0; c < kC; ++c)
          v[c] = (ch < tk.nch && k0 + c < K) ? base[(size_t)ch * K + k0 + c] : 0u;
        uint64_t sl = 0, d = 0, sle = 0;  // at the chunk's start: sum c^2, sum t c; at its end: sum c^2
#pragma unroll
        for (uint32_t c = 0; c < kC; ++c) {
          const uint32_t inc = wave_incl_scan(v[c]);
          const uint32_t before = carry[c] + inc - v[c], after = carry[c] + inc;
          if (ch < tk.nch && k0 + c < K) base[(size_t)ch * K + k0 + c] = before;
          sl += (uint64_t)before * before;
          d += (uint64_t)t[c] * before;
          sle += (uint64_t)after * after;
          carry[c] += readlane(inc, kWave - 1);
        }
        if (k0 == 0 && ch < tk.nch) {
          double ub = -1.0;  // no bound: the chunk is kept
          if (prune) {
            const uint64_t SRs = T2 + sl - 2 * d;
            if (ch + 1 < tk.nch) {
              const uint32_t p0 = ch * a.Cs, p1 = p0 + a.Cs;  // p1 < n: a later chunk exists
              ub = (double)sle / (double)(p0 + 1u) + (double)SRs / (double)(n - p1);
            }
            if (bounds && ch >= 1) {
              // the boundary before the chunk: positions p, p + 1 < n
              const uint32_t p = ch * a.Cs - 1u;
              const uint32_t s0 = sample_of(row[p]), s1 = sample_of(row[p + 1u]);
              const uint32_t w0 = kind ? code_at(cb, kind, s0) : __float_as_uint(x[s0]);
              const uint32_t w1 = kind ? code_at(cb, kind, s1) : __float_as_uint(x[s1]);
              if (kind ? w0 != w1 : gap_ok(__uint_as_float(w0), __uint_as_float(w1))) {
                const uint64_t nL = (uint64_t)ch * a.Cs, nR = (uint64_t)n - nL;
                known = fmax(known, (double)(sl * nR + SRs * nL) / (double)(nL * nR));
              }
            }
          }
          // without the boundaries the known score is final: mark now
          const bool skip = !bounds && ub >= 0.0 && ub * (1.0 + 0x1p-30) < known * (1.0 - 0x1p-30);
          if (skip) skipped += (unsigned long long)a.Cs * (sizeof(IdxT) + sizeof(LabT) + 4);
          ScanRes r;
          r.num = bounds ? (unsigned long long)__double_as_longlong(ub) : 0ull;
          r.den = 1;
          r.pos = 0;
          r.feat = f;
          r.ok = 0;
          r.pad = skip ? kScanSkipped : 0u;
          rs[ch] = r;
        }
      }
    }
    if (bounds) {
      // the best known score over the lanes, rounded down
#pragma unroll
      for (uint32_t s = kWave / 2; s > 0; s >>= 1) known = fmax(known, __shfl_xor(known, (int)s, kWave));
      const double thr = known * (1.0 - 0x1p-30);
      for (uint32_t c0 = 0; c0 < tk.nch; c0 += kWave) {
        const uint32_t ch = c0 + lane;
        if (ch >= tk.nch) break;
        // this lane stored the entry above
        const double ub = __longlong_as_double((long long)rs[ch].num);
        const bool skip = ub >= 0.0 && ub * (1.0 + 0x1p-30) < thr;
        if (skip) skipped += (unsigned long long)a.Cs * (sizeof(IdxT) + sizeof(LabT) + 4);
        ScanRes r;
        r.num = 0;
        r.den = 1;
        r.pos = 0;
        r.feat = f;
        r.ok = 0;
        r.pad = skip ? kScanSkipped : 0u;
        rs[ch] = r;
      }
    }
  }
  flush_bytes(a, kBytesScanSkip, wave_sum64(skipped));
}

// ------------------------------------------------------------------------
// k_scan: Gini scan of the task's chunk in each of its features.  A wave
// keeps kScanU 64-position tiles in flight: their index loads, then their
// label and value gathers, are issued together before the tiles are scanned
// in order.
// ------------------------------------------------------------------------
#ifndef CTA_SCAN_U
#define CTA_SCAN_U 3
#endif
#ifndef CTA_SCAN_WAVES
#define CTA_SCAN_WAVES 6
#endif
constexpr int kScanU = CTA_SCAN_U;
// Label ranks in the split scan: with the class counts in registers (lane k
// holds class k), one ballot per distinct label of a tile; with them in LDS
// (more classes), one ballot per label bit.
#ifndef CTA_SCAN_REGS_K
#define CTA_SCAN_REGS_K 4
#endif
#ifndef CTA_RANK_BITS
#define CTA_RANK_BITS 1
#endif
constexpr uint32_t kScanRegsK = CTA_SCAN_REGS_K;
constexpr bool kRankByBits = CTA_RANK_BITS != 0;

// Gini scan of positions [p0, p1) of a node's sorted row (n samples, feature
// f): the best candidate position of the range, in every lane.  SL and D are
// the running sums S_L and sum_k tot_k c_k at p0; cv (K <= 64 with regs: lane
// k holds class k) or carry[] (LDS) the class counts before p0; tv / tot the
// node's class totals, T2 their sum of squares; chunk_next the value at p1
// when p1 < n.  vat(s) is sample s's value (code or float bits) and
// cand(v, w) the candidate test between adjacent values.  kScanU tiles of 64
// positions are in flight: their index loads, then their label and value
// gathers, are issued together before the tiles are scanned in order.
template <bool regs, class IdxT, class LabT, class VAt, class CandF>
__device__ __forceinline__ Cand scan_positions(const FitArgs<IdxT, LabT> &a, const IdxT *row, uint32_t p0,
                                               uint32_t p1, uint32_t n, uint32_t f, uint32_t cv, uint32_t tv,
                                               uint32_t *carry, const uint32_t *tot, uint64_t SL, uint64_t D,
                                               uint64_t T2, uint32_t chunk_next, VAt vat, CandF cand) {
  const uint32_t lane = lane_id();
  Cand lb = {0, 1, 0, (int32_t)f, 0};
  for (uint32_t pb = p0; pb < p1; pb += kWave * kScanU) {
    // kScanU tiles in flight: index loads, then label and value gathers,
    // all unmasked (positions past p1 - 1 read the last entry; the scan
    // below masks them) so that each group issues without waits between
    uint32_t sid[kScanU], lab[kScanU], val[kScanU];
#pragma unroll
    for (int u = 0; u < kScanU; ++u) sid[u] = (uint32_t)row[min(pb + (uint32_t)u * kWave + lane, p1 - 1u)];
    const uint32_t pe = pb + kWave * kScanU;  // first position after the batch
    const uint32_t e_next = (uint32_t)row[min(pe, p1 - 1u)];  // one address for the wave
#pragma unroll
    for (int u = 0; u < kScanU; ++u) {
      lab[u] = a.label((IdxT)sid[u]);
      val[u] = vat(sample_of((IdxT)sid[u]));
    }
    const uint32_t batch_next = vat(sample_of((IdxT)e_next));  // used only when pe < p1
#pragma unroll
    for (int u = 0; u < kScanU; ++u) {
      const uint32_t p = pb + (uint32_t)u * kWave + lane;
      const bool valid = p < p1;
      if (u > 0 && !ballot(valid)) break;  // the first tile holds position pb < p1
      // rank of this lane among earlier lanes with the same label (one
      // ballot per distinct label of the tile), and its class total
      uint32_t cprev = 0, dD = 0;
      uint64_t rem = ballot(valid);
      if constexpr (regs) {
        while (rem) {
          const uint32_t leader = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
          const uint32_t L = readlane(lab[u], leader);
          const uint64_t mm = ballot(valid && lab[u] == L);
          const uint32_t b0 = readlane(cv, L), tL = readlane(tv, L);
          if (valid && lab[u] == L) {
            cprev = b0 + mbcnt(mm);
            dD = tL;
          }
          if (lane == L) cv += (uint32_t)__popcll(mm);
          rem &= ~mm;
        }
      } else if (kRankByBits) {
        // the lanes holding this lane's label are those that agree with it
        // on every label bit (one ballot per bit, whatever the number of
        // labels in the tile); the last of them stores the label's new count
        const uint32_t nbits = a.K > 1 ? 32u - (uint32_t)__builtin_clz((uint32_t)a.K - 1u) : 1u;
        uint64_t same = rem;
        for (uint32_t b = 0; b < nbits; ++b) {
          const uint32_t bit = (lab[u] >> b) & 1u;
          const uint64_t mb = ballot(bit != 0u);
          same &= bit ? mb : ~mb;
        }
        if (valid) {
          cprev = carry[lab[u]] + mbcnt(same);
          dD = tot[lab[u]];
        }
        wave_sync();
        if (valid && (same >> lane) == 1ull) carry[lab[u]] = cprev + 1u;
        wave_sync();
      } else {
        while (rem) {
          const uint32_t leader = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
          const uint32_t L = readlane(lab[u], leader);
          const uint64_t mm = ballot(valid && lab[u] == L);
          const uint32_t b0 = carry[L];
          if (valid && lab[u] == L) cprev = b0 + mbcnt(mm);
          wave_sync();
          if (lane == 0) carry[L] = b0 + (uint32_t)__popcll(mm);
          wave_sync();
          rem &= ~mm;
        }
        if (valid) dD = tot[lab[u]];
      }
      const uint32_t dS = valid ? 2u * cprev + 1u : 0u;
      const uint32_t iS = wave_incl_scan(dS);
      const uint32_t iD = wave_incl_scan(dD);
      const uint64_t SLp = SL + iS, Dp = D + iD;
      SL += readlane(iS, kWave - 1);
      D += readlane(iD, kWave - 1);
      // value at p + 1: the next lane, the next tile's lane 0, the next
      // batch, or the next chunk
      uint32_t nxt = wave_next(val[u]);
      if (u + 1 < kScanU) {
        const uint32_t first_next = readlane(val[u + 1 < kScanU ? u + 1 : u], 0);
        if (lane == kWave - 1) nxt = first_next;
      } else if (lane == kWave - 1) {
        nxt = batch_next;
      }
      if (p + 1 == p1) nxt = chunk_next;
      const bool elig = valid && (p + 1 < n) && cand(val[u], nxt);
      if (elig) {
        const uint64_t nL = (uint64_t)p + 1, nR = (uint64_t)n - nL;
        const uint64_t SR = T2 + SLp - 2 * Dp;
        Cand c = {SLp * nR + SR * nL, nL * nR, p, (int32_t)f, 1u};
        if (!lb.ok || q_cmp(c.num, c.den, lb.num, lb.den) > 0) lb = c;
      }
    }
  }
  return wave_best(lb);
}

// kRegs: the class counts of a wave in lane registers (lane k holds class
// k; at most kScanRegsK classes), else in LDS.  The host launches the one
// instance that fits K.
template <class IdxT, class LabT, bool kRegs>
// Six waves per SIMD (a few registers spilled) beat five without spills.
__global__ __launch_bounds__(kBlock, CTA_SCAN_WAVES) void k_scan(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  uint32_t *carry = smem + smem_wave_offset(a.K);
  uint32_t *tot = carry + K;
  // workgroup b takes the task list of class b % kXcd (its XCD's features)
  const uint32_t cls = blockIdx.x % kXcd;
  const uint32_t cn = lane < kXcd ? a.ctl->n_cls[lane] : 0u;
  const uint32_t cbase = wave_incl_scan(cn) - cn;
  const uint32_t lo = readlane(cbase, cls), hi = min(lo + readlane(cn, cls), a.cap_scan);
  const uint32_t nwc = (gridDim.x / kXcd) * kWavesPerBlock;
  for (uint32_t item = lo + (blockIdx.x / kXcd) * kWavesPerBlock + threadIdx.x / kWave; item < hi; item += nwc) {
    const Task tk = a.scan_tasks[item];
    const uint32_t i = tk.node;
    const FNode nd = a.front[i];
    const NAux ax = a.aux[i];
    // a chunk k_scan_carry ruled out keeps the result entry it wrote
    if (tk.nch > 1 && a.res[ax.scan_base + (size_t)(tk.first_count & 0xFFFFu) * tk.nch + tk.chunk].pad == kScanSkipped)
      continue;
    const uint32_t n = nd.n;
    // class totals: lane k holds class k (kRegs), else in LDS
    constexpr bool regs = kRegs;
    uint64_t t2 = 0;
    uint32_t tv = 0;
    for (uint32_t k = lane; k < K; k += kWave) {
      const uint32_t v = a.cnt[(size_t)i * K + k];
      if (regs) tv = v;
      else tot[k] = v;
      t2 += (uint64_t)v * v;
    }
    const uint64_t T2 = uniform64(wave_sum64(t2));
    const uint32_t first = tk.first_count & 0xFFFFu, count = tk.first_count >> 16;
    const uint32_t p0 = tk.chunk * a.Cs, p1 = (tk.nch > 1) ? min(n, p0 + a.Cs) : n;
    Cand best = {0, 1, 0, 0, 0};
    for (uint32_t jj = 0; jj < count; ++jj) {
      const uint32_t j = first + jj;
      const uint32_t f = (uint32_t)a.feats[(size_t)i * a.m + j];
      const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
      const float *x = a.X_T + (size_t)f * a.N;
      // the feature's value code: its distinct-value index from a 1/8/16-bit
      // column (a candidate exactly where the index changes), or the float
      // bits (a candidate where the gap test passes)
      const uint64_t fc = a.fcode[f];
      const uint32_t kind = uniform(fc_kind(fc));
      const uint8_t *cb = a.codes + fc_off(fc);
      auto vat = [&](uint32_t s) -> uint32_t { return kind ? code_at(cb, kind, s) : __float_as_uint(x[s]); };
      auto cand = [&](uint32_t v, uint32_t w) -> bool {
        return kind ? v != w : gap_ok(__uint_as_float(v), __uint_as_float(w));
      };
      auto differ = [](uint32_t v, uint32_t w) -> bool { return v != w; };
      if (kind == 1u && tk.nch == 1 && regs && n > kWave * kScanU) {
        // Two-valued feature (the node holds both values: the selection
        // dropped constant features): its one candidate is the position
        // before the first 1.  Find it by a 64-way search over the row, then
        // count the labels of the shorter side; the Gini terms are the same
        // integers the full scan forms at that position.
        uint32_t lo = 0, hi = n - 1u;  // code 0 at lo, code 1 at hi
        while (hi - lo > 1u) {
          const uint32_t span = hi - lo;
          uint32_t pr = lo + (uint32_t)(((uint64_t)span * (lane + 1u)) / (kWave + 1u));
          pr = min(max(pr, lo + 1u), hi - 1u);
          // probes rise with the lane, so the lanes reading 0 come first
          const uint64_t z0 = ballot(code_at(cb, 1u, sample_of(row[pr])) == 0u);
          if (z0) lo = readlane(pr, 63u - (uint32_t)__clzll((unsigned long long)z0));
          if (~z0) hi = readlane(pr, (uint32_t)__ffsll((unsigned long long)~z0) - 1u);
        }
        const uint32_t z = hi;  // samples with code 0
        const bool left = z <= n - z;
        const uint32_t q0 = left ? 0u : z, q1 = left ? z : n;
        uint32_t cnt = 0;  // lane k: samples of class k in [q0, q1)
        for (uint32_t pb = q0; pb < q1; pb += kWave * kScanU) {
          uint32_t l[kScanU];
#pragma unroll
          for (int u = 0; u < kScanU; ++u) {
            const uint32_t p = pb + (uint32_t)u * kWave + lane;
            l[u] = (p < q1) ? a.label(row[p]) : 0xFFFFFFFFu;
          }
#pragma unroll
          for (int u = 0; u < kScanU; ++u) {
            uint64_t rem = ballot(l[u] != 0xFFFFFFFFu);
            while (rem) {
              const uint32_t L = readlane(l[u], (uint32_t)__ffsll((unsigned long long)rem) - 1u);
              const uint64_t mm = ballot(l[u] == L);
              if (lane == L) cnt += (uint32_t)__popcll(mm);
              rem &= ~mm;
            }
          }
        }
        const uint32_t cl = left ? cnt : tv - cnt;  // left-child count of class `lane`
        const uint64_t SLz = uniform64(wave_sum64((uint64_t)cl * cl));
        const uint64_t Dz = uniform64(wave_sum64((uint64_t)tv * cl));
        const uint64_t nL = z, nR = (uint64_t)n - z;
        const uint64_t SR = T2 + SLz - 2 * Dz;
        const Cand c = {SLz * nR + SR * nL, nL * nR, z - 1u, (int32_t)f, 1u};
        if (cand_better(c, best)) best = c;
        continue;
      }
      // Values never decrease along the row, so when the first and the last
      // value a chunk compares (positions p0 and min(p1, n - 1)) fail the
      // candidate test, every adjacent pair between them fails it too: such
      // a chunk (all but one chunk of a two-valued feature) is skipped.
      const uint32_t pe = min(p1, n - 1u);
      const uint32_t vpe = (p1 < n || tk.nch > 1) ? vat(sample_of(row[pe])) : 0u;
      if (tk.nch > 1 && (p0 >= pe || !cand(vat(sample_of(row[p0])), vpe))) continue;
      // chunk carry and the running sums at the chunk start; with K <= 64
      // lane k carries class k's count in cv
      wave_sync();
      uint64_t sl = 0, dd = 0;
      uint32_t cv = 0;
      const uint32_t *hc = (tk.nch > 1) ? a.hist + ax.hist_base + ((size_t)j * tk.nch + tk.chunk) * K : nullptr;
      for (uint32_t k = lane; k < K; k += kWave) {
        const uint32_t c = hc ? hc[k] : 0u;
        const uint32_t t = regs ? tv : tot[k];
        if (regs) cv = c;
        else carry[k] = c;
        sl += (uint64_t)c * c;
        dd += (uint64_t)t * c;
      }
      uint64_t SL = uniform64(wave_sum64(sl));
      uint64_t D = uniform64(wave_sum64(dd));
      wave_sync();
      // value just past the chunk end (the next chunk's first sample)
      const uint32_t chunk_next = (p1 < n) ? vpe : 0u;
      // one instance per column kind, so the position loop's gathers carry
      // no branches on the kind
      auto scan = [&](auto at, auto test) -> Cand {
        return scan_positions<kRegs>(a, row, p0, p1, n, f, cv, tv, carry, tot, SL, D, T2, chunk_next, at, test);
      };
      Cand lb;
      if (kind == 2u)
        lb = scan([&](uint32_t s) { return code_at_k<2>(cb, s); }, differ);
      else if (kind == 3u)
        lb = scan([&](uint32_t s) { return code_at_k<3>(cb, s); }, differ);
      else if (kind == 1u)
        lb = scan([&](uint32_t s) { return code_at_k<1>(cb, s); }, differ);
      else
        lb = scan([&](uint32_t s) { return __float_as_uint(x[s]); }, cand);
      if (cand_better(lb, best)) best = lb;
    }
    if (lane == 0) {
      ScanRes r;
      r.num = best.num;
      r.den = best.den;
      r.pos = best.pos;
      r.feat = best.feat;
      r.ok = best.ok;
      r.pad = 0;
      // results in node order: the node's task index t as emitted
      const uint32_t first = tk.first_count & 0xFFFFu;
      a.res[ax.scan_base + (tk.nch > 1 ? first * tk.nch + tk.chunk : first / ax.fpt)] = r;
    }
  }
}

// ------------------------------------------------------------------------
// Reshuffle geometry.  Rows of a node with at least kBigN samples are moved
// by whole 1,024-thread workgroups (one task per row, kReshTile-element
// tiles); smaller nodes by single waves (one task per group of rows holding
// about Cr elements).
// ------------------------------------------------------------------------
#ifndef CTA_BIG_N
#define CTA_BIG_N 4096
#endif
constexpr uint32_t kBigN = CTA_BIG_N;
// A node with at least kBigMaxRows rows gives the level enough waves with
// one wave per group of rows, which needs no workgroup barriers: its rows
// never take big tasks.
#ifndef CTA_BIG_MAX_ROWS
#define CTA_BIG_MAX_ROWS 1024
#endif
constexpr uint32_t kBigMaxRows = CTA_BIG_MAX_ROWS;
constexpr int kReshThreads = 1024;
constexpr int kReshWaves = kReshThreads / kWave;  // 16
#ifndef CTA_RESH_SMALL_E
#define CTA_RESH_SMALL_E 4
#endif
constexpr int kReshE = CTA_RESH_SMALL_E;  // elements per lane in flight in a small task
// Big tasks: elements per thread per tile, and whether the next tile's
// loads are issued into a second register set before this tile's stores.
#ifndef CTA_RESH_BIG_E
#define CTA_RESH_BIG_E 4
#endif
#ifndef CTA_RESH_PIPE
#define CTA_RESH_PIPE 1
#endif
constexpr int kBigE = CTA_RESH_BIG_E;
constexpr bool kBigPipe = CTA_RESH_PIPE != 0;
constexpr uint32_t kReshTile = kReshThreads * kBigE;
constexpr uint32_t kSlots = kBigE * kReshWaves;  // (sub-tile, wave) counts per tile
constexpr uint32_t kSpl = kSlots / kWave;        // of them per lane in the tile prefix
static_assert(kSlots % kWave == 0 && kReshWaves % kSpl == 0, "the tile prefix takes whole slots per lane");
#ifndef CTA_WG_ELEMS_MIN_LOG2
#define CTA_WG_ELEMS_MIN_LOG2 16
#endif
constexpr uint64_t kWgElemsMax = 1ull << 19, kWgElemsMin = 1ull << CTA_WG_ELEMS_MIN_LOG2;

// Label of the majority class (lowest label on ties).
__device__ uint32_t wave_majority(const uint32_t *c, uint32_t K) {
  const uint32_t lane = lane_id();
  uint32_t bv = 0, bk = 0xFFFFFFFFu;
  for (uint32_t k = lane; k < K; k += kWave)
    if (bk == 0xFFFFFFFFu || c[k] > bv) {
      bv = c[k];
      bk = k;
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t ov = __shfl_xor(bv, o, kWave), ok = __shfl_xor(bk, o, kWave);
    if (ok != 0xFFFFFFFFu && (bk == 0xFFFFFFFFu || ov > bv || (ov == bv && ok < bk))) {
      bv = ov;
      bk = ok;
    }
  }
  return bk;
}

// Reshuffle task shape of a node of n samples with a live child; fills
// v[2..4] (big tasks, small tasks, big tiles) of its children-scan counts.
__device__ __forceinline__ void resh_shape(uint32_t n, uint32_t F, uint32_t Cr, NAux &ax, uint32_t *v) {
  const bool big = n >= kBigN && F < kBigMaxRows;
  const uint32_t rpt = big ? 1u : max(1u, Cr / n);
  const uint32_t ntask = big ? F : ceil_div(F, rpt);
  ax.resh_big = big ? 1u : 0u;
  ax.rpt = rpt;
  ax.resh_ntask = ntask;
  v[2] = big ? F : 0u;
  v[3] = big ? 0u : ntask;
  v[4] = big ? F * ceil_div(n, kReshTile) : 0u;
}

// ------------------------------------------------------------------------
// k_decide: best split of each node over its tasks, threshold, class counts
// of the left child and the node's bitmap-fill tasks; for a split node, the
// leaf tests of both children (random_tree.py:123-163) and its
// children-scan counts: child ids, next-frontier slots, reshuffle tasks.
// ------------------------------------------------------------------------

// Class counts of positions [0, p] of row f of node i into cntL.  The chunk
// holding p is recounted from whichever of its ends is nearer: forwards
// from the chunk's carried counts, or backwards (subtracting) from the next
// chunk's carried counts (the node's counts for the last chunk).  Carries
// exist for multi-chunk scans only; a single-chunk node is one chunk
// starting from zero.
template <class IdxT, class LabT>
__device__ void count_left(const FitArgs<IdxT, LabT> &a, uint32_t i, const FNode &nd, const NAux &ax, uint32_t f,
                           uint32_t p, uint32_t *h) {
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K, n = nd.n, nch = ax.nch_s;
  const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
  const uint32_t c = (nch > 1) ? p / a.Cs : 0u;
  const uint32_t c0 = c * a.Cs, c1 = (nch > 1) ? min(n, c0 + a.Cs) : n;
  const uint32_t *before = nullptr;                 // counts before c0 (nullptr: zero)
  const uint32_t *after = a.cnt + (size_t)i * K;    // counts before c1
  if (nch > 1) {
    // slot of f among the node's evaluated features
    uint32_t j = 0;
    for (uint32_t j0 = 0; j0 < ax.nf; j0 += kWave) {
      const uint32_t jj = j0 + lane;
      const uint64_t mk = ballot(jj < ax.nf && (uint32_t)a.feats[(size_t)i * a.m + jj] == f);
      if (mk) {
        j = j0 + (uint32_t)__ffsll((unsigned long long)mk) - 1u;
        break;
      }
    }
    const uint32_t *hb = a.hist + ax.hist_base + (size_t)j * nch * K;
    before = hb + (size_t)c * K;
    if (c + 1 < nch) after = hb + (size_t)(c + 1) * K;
  }
  const bool fwd = (p + 1 - c0) <= (c1 - (p + 1));
  wave_sync();
  for (uint32_t k = lane; k < K; k += kWave) h[k] = fwd ? (before ? before[k] : 0u) : after[k];
  wave_sync();
  const uint32_t r0 = fwd ? c0 : p + 1, r1 = fwd ? p + 1 : c1;
  const uint32_t delta = fwd ? 1u : 0xFFFFFFFFu;
  for (uint32_t pb = r0; pb < r1; pb += 4 * kWave) {
    uint32_t sid[4];
    load_row_batch(row, pb, r1, sid);
    uint32_t l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) l[u] = a.label((IdxT)sid[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (pb + (uint32_t)u * kWave + lane >= r1) l[u] = 0xFFFFFFFFu;  // past the range
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (l[u] != 0xFFFFFFFFu) atomicAdd(&h[l[u]], delta);
  }
  wave_sync();
  for (uint32_t k = lane; k < K; k += kWave) a.cntL[(size_t)i * K + k] = h[k];
  wave_sync();
}

// What a child of a node at `depth` becomes, from its count of present
// classes nz and its size nc: 0 live (split further), 1 pure leaf, 2 leaf
// labelled by its first sample (fewer than mss samples), 3 majority leaf
// (depth cap).
template <class IdxT, class LabT>
__device__ __forceinline__ uint32_t child_kind(const FitArgs<IdxT, LabT> &a, uint32_t nz, uint32_t nc,
                                               uint32_t depth) {
  if (nz <= 1) return 1u;
  if ((int64_t)nc < (int64_t)a.mss) return 2u;
  if (a.max_depth > 0 && (int64_t)depth + 1 >= (int64_t)a.max_depth) return 3u;
  return 0u;
}

template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_decide(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  uint32_t *h = smem + smem_wave_offset(a.K);
  unsigned long long bytes = 0, bytes_resh = 0, bytes_resh_small = 0;
  WAVE_LOOP(a.ctl->q) {
    const uint32_t i = item;
    const FNode nd = a.front[i];
    NAux ax = a.aux[i];
    Cand best = {0, 1, 0, 0, 0};
    for (uint32_t t = lane; t < ax.scan_ntask; t += kWave) {
      const ScanRes r = a.res[ax.scan_base + t];
      Cand c = {r.num, r.den, r.pos, r.feat, r.ok};
      if (cand_better(c, best)) best = c;
    }
    best = wave_best(best);
    // the evaluated row-free features' best candidate, under the same order
    const Cand rlcand = {ax.rl_num, ax.rl_den, ax.rl_pos, ax.rl_feat, (a.Frl && ax.rl_feat >= 0) ? 1u : 0u};
    const bool rowfree = cand_better(rlcand, best);
    if (rowfree) best = rlcand;
    const uint32_t id = nd.nid;
    uint32_t nfill = 0;
    if (!best.ok) {
      // no candidate position: leaf labelled by the first sample of the
      // segment in feature 0's row (random_tree.py:628-634, 50-63)
      if (lane == 0) {
        const IdxT *row0 = rows_of(a, nd.src, nd.tree, 0) + nd.start;
        a.arena.feat[id] = -1;
        a.arena.thr[id] = 0.f;
        a.arena.left[id] = kLeaf;
        a.arena.right[id] = kLeaf;
        a.arena.value[id] = a.label(row0[0]);
      }
      ax.best_f = -1;
      ax.flags = 0;
    } else if (rowfree) {
      // lower value left: per-class left counts are totals minus the counts
      // holding the higher value; the side bits come from the value bits
      // over the node's whole feature-0 row
      const uint32_t f = (uint32_t)best.feat, p = best.pos;
      const uint32_t K = (uint32_t)a.K;
      const uint32_t *oc = a.rlc + ((size_t)i * a.Frl + (uint32_t)a.rl_of[f]) * K;
      for (uint32_t k = lane; k < K; k += kWave) a.cntL[(size_t)i * K + k] = a.cnt[(size_t)i * K + k] - oc[k];
      nfill = nd.n;
      if (lane == 0) {
        a.arena.feat[id] = (int32_t)f;
        a.arena.thr[id] = a.rl_thr[a.rl_of[f]];
        a.arena.value[id] = 0;
        bytes += (unsigned long long)nd.n * (sizeof(IdxT) + 1);  // fill: n (b + 1)
      }
      ax.best_f = (int32_t)f;
      ax.best_p = p;
      ax.nL = p + 1;
      ax.flags = kSplit | (2 * ax.nL > nd.n ? kInvert : 0u);
      ax.dst = (nd.src == 2) ? 1u : 2u;
    } else {
      const uint32_t f = (uint32_t)best.feat, p = best.pos;
      count_left(a, i, nd, ax, f, p, h);
      nfill = min(p + 1, nd.n - (p + 1));  // the smaller side's positions; chunks below
      if (lane == 0) {
        const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
        const float *x = a.X_T + (size_t)f * a.N;
        // threshold midway in float, no contraction (bfs_module.cu:399-400)
        const float thr = __fmul_rn(__fadd_rn(x[sample_of(row[p])], x[sample_of(row[p + 1])]), 0.5f);
        a.arena.feat[id] = (int32_t)f;
        a.arena.thr[id] = thr;
        a.arena.value[id] = 0;
        bytes += (unsigned long long)nfill * (sizeof(IdxT) + 1);  // fill: the smaller side (b + 1)
      }
      ax.best_f = (int32_t)f;
      ax.best_p = p;
      ax.nL = p + 1;
      ax.flags = kSplit | (2 * ax.nL > nd.n ? kInvert : 0u);
      ax.dst = (nd.src == 2) ? 1u : 2u;
    }
    uint32_t v[kChFields] = {0u, 0u, 0u, 0u, 0u, 0u};  // children-scan counts
    if (ax.flags & kSplit) {
      // both children: class counts (lane k: class k), how many classes
      // each holds and the single present one (random_tree.py:123-163)
      const uint32_t K = (uint32_t)a.K, nL = ax.nL, nR = nd.n - ax.nL;
      uint32_t *cl = h, *cr = h + K;
      uint32_t nzl = 0, nzr = 0, onlyl = 0, onlyr = 0;
      wave_sync();
      for (uint32_t k = lane; k < K; k += kWave) {
        const uint32_t l = a.cntL[(size_t)i * K + k];  // written by this lane above
        const uint32_t r = a.cnt[(size_t)i * K + k] - l;
        cl[k] = l;
        cr[k] = r;
        if (l) {
          ++nzl;
          onlyl = k;
        }
        if (r) {
          ++nzr;
          onlyr = k;
        }
      }
      nzl = uniform(wave_sum32(nzl));
      nzr = uniform(wave_sum32(nzr));
      // the single present class (max over lanes: absent lanes hold 0)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        onlyl = max(onlyl, (uint32_t)__shfl_xor(onlyl, o, kWave));
        onlyr = max(onlyr, (uint32_t)__shfl_xor(onlyr, o, kWave));
      }
      wave_sync();
      const uint32_t kl = child_kind(a, nzl, nL, nd.depth), kr = child_kind(a, nzr, nR, nd.depth);
      uint32_t labl = onlyl, labr = onlyr;
      if (kl == 3) labl = wave_majority(cl, K);
      if (kr == 3) labr = wave_majority(cr, K);
      // a leaf labelled by its first sample needs the side bits: k_children_emit
      // looks it up after the fill
      if (kl == 2) labl = kFirstSample;
      if (kr == 2) labr = kFirstSample;
      ax.lab[0] = kl ? labl : kLeaf;
      ax.lab[1] = kr ? labr : kLeaf;
      if (kl == 0) ax.flags |= kLiveL;
      if (kr == 0) ax.flags |= kLiveR;
      v[0] = 2;
      // a live child of at most sub_max samples goes to the small-subtree
      // queue, a larger one to the next frontier
      const uint32_t subl = (kl == 0 && nL <= a.sub_max) ? 1u : 0u, subr = (kr == 0 && nR <= a.sub_max) ? 1u : 0u;
      v[1] = (kl == 0 ? 1u : 0u) + (kr == 0 ? 1u : 0u) - subl - subr;
      v[5] = subl + subr;
      ax.resh_ntask = 0;
      if (kl == 0 || kr == 0) {
        resh_shape(nd.n, a.Fr, a.Cr, ax, v);
        // every row read and written, the mark gathered per element: Fr n (2b + 1)
        if (lane == 0) {
          const unsigned long long b = (unsigned long long)a.Fr * nd.n * (2 * sizeof(IdxT) + 1);
          if (ax.resh_big) bytes_resh += b;
          else bytes_resh_small += b;
        }
      }
      // a live child that stays in the levels needs row-free counts: the
      // fill tasks count the smaller side's samples, the ones they mark,
      // into rl_small (from zero)
      if (a.Frl && ((kl == 0 && nL > a.sub_max) || (kr == 0 && nR > a.sub_max))) {
        ax.flags |= kCountRl;
        // (a node of one fill task gets its counts stored whole by that task)
        if (nfill > a.Cf)
          for (uint32_t w = lane; w < a.Frl * K; w += kWave) a.rl_small[(size_t)i * a.Frl * K + w] = 0u;
        // signature words of the smaller side
        if (lane == 0) bytes += (unsigned long long)min(nL, nR) * 4 * a.Wsig;
      }
    }
    if (lane == 0) {
      uint32_t *o = a.vals + (size_t)i * kChFields;
#pragma unroll
      for (int f = 0; f < kChFields; ++f) o[f] = v[f];
    }
    nfill = ceil_div(nfill, a.Cf);
    // the node's fill slots, one per Cf positions of its rows: the side bits
    // of the smaller side and its row-free counts come from the first nfill;
    // a node that does not split leaves them empty
    for (uint32_t c = lane; c < ax.nch_f; c += kWave)
      a.fill_tasks[ax.fill_base + c] = (ax.flags & kSplit) ? Task{i, 0u, c, nfill} : Task{kLeaf, 0u, 0u, 0u};
    if (lane == 0) a.aux[i] = ax;
  }
  flush_bytes(a, kBytesFill, bytes);
  flush_bytes(a, kBytesResh, bytes_resh);
  flush_bytes(a, kBytesReshSmall, bytes_resh_small);
}

// ------------------------------------------------------------------------
// Side bitmaps: bit s of tree t is set when sample s goes to the smaller
// child of its node: the left one, or the right one when the node is flagged
// kInvert (dfs_module.cu:210-221, bfs_module.cu:238-264 mark the left
// child; marking the smaller one halves the marks on average).
//
// k_fill_lds: workgroup (t, j) of J per tree builds the bits of its share of
// tree t's fill tasks in an LDS copy of the bitmap and writes it to the
// tree's bitmap, which k_select cleared at the start of the level: the
// whole copy when the tree has one fill workgroup, else its non-zero words
// by atomic OR.
// ------------------------------------------------------------------------
constexpr int kFillThreads = 1024;
constexpr int kFillWaves = kFillThreads / kWave;

// Adds the signatures of the entries e[u] with take[u] set to the wave's
// row-free counters: cnt[j K + c] counts the samples of class c holding the
// higher value of row-free feature j.  The gathers of all U entries are
// issued before the LDS adds.
template <int U, class IdxT, class LabT>
__device__ __forceinline__ void rl_add(const FitArgs<IdxT, LabT> &a, const uint32_t (&e)[U], const bool (&take)[U],
                                       uint32_t *cnt) {
  const uint32_t K = (uint32_t)a.K, Ws = a.Wsig;
  uint32_t w0[U], w1[U], lab[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    w0[u] = w1[u] = lab[u] = 0;
    if (take[u]) {
      const uint32_t s = sample_of((IdxT)e[u]);
      lab[u] = a.label((IdxT)e[u]);
      w0[u] = a.sig[(size_t)s * Ws];
      if (Ws > 1) w1[u] = a.sig[(size_t)s * Ws + 1];
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (!take[u]) continue;
    uint32_t *c = cnt + lab[u];
    for (uint32_t w = 0; w < Ws; ++w) {
      uint32_t word = w == 0 ? w0[u] : w == 1 ? w1[u] : a.sig[(size_t)sample_of((IdxT)e[u]) * Ws + w];
      while (word) {
        atomicAdd(&c[(w * 32u + (uint32_t)__builtin_ctz(word)) * K], 1u);
        word &= word - 1u;
      }
    }
  }
}

// positions per lane in flight in a fill task
#ifndef CTA_FILL_U
#define CTA_FILL_U 8
#endif
constexpr int kFillU = CTA_FILL_U;

// Adds the wave's row-free counters (node `node`'s) to rl_small and clears
// them.  `whole`: this wave counted the node's only fill task, so its
// counters are the node's counts and are stored outright (rl_small was not
// cleared for such a node); otherwise the non-zero ones are added.
template <class IdxT, class LabT>
__device__ __forceinline__ void rl_flush(const FitArgs<IdxT, LabT> &a, uint32_t node, uint32_t *cnt, bool whole) {
  const uint32_t lane = lane_id(), FK = a.Frl * (uint32_t)a.K;
  wave_sync();
  uint32_t *dst = a.rl_small + (size_t)node * FK;
  if (whole) {
    for (uint32_t w = lane; w < FK; w += kWave) {
      dst[w] = cnt[w];
      cnt[w] = 0;
    }
  } else {
    for (uint32_t w = lane; w < FK; w += kWave)
      if (cnt[w]) {
        atomicAdd(&dst[w], cnt[w]);
        cnt[w] = 0;
      }
  }
  wave_sync();
}

// Clears a wave's row-free counters (before its first fill task).
__device__ __forceinline__ void rl_zero(uint32_t *cnt, uint32_t FK) {
  for (uint32_t w = lane_id(); w < FK; w += kWave) cnt[w] = 0;
  wave_sync();
}

// One fill task: the marks of the smaller side's samples among Cf positions
// of a split node and, when the node is flagged kCountRl, their row-free
// counts, added to the wave's counters cnt.  The counters (zero at the
// wave's start) hold node `cur`'s counts: a task of another counting node
// flushes them to rl_small first; the wave flushes the last node's at its
// end.
template <class IdxT, class LabT>
__device__ __forceinline__ void fill_task_bits(const FitArgs<IdxT, LabT> &a, const Task &tk, uint32_t *bits,
                                               uint32_t *cnt, uint32_t &cur, bool &cur_whole) {
  const uint32_t lane = lane_id();
  const FNode nd = a.front[tk.node];
  const NAux ax = a.aux[tk.node];
  const int32_t jr = a.Frl ? a.rl_of[ax.best_f] : -1;
  const bool count = (ax.flags & kCountRl) != 0u;
  const bool small_left = ax.nL <= nd.n - ax.nL;
  if (count && tk.node != cur) {
    if (cur != kLeaf) rl_flush(a, cur, cnt, cur_whole);
    cur = tk.node;
    cur_whole = tk.nch == 1;
  }
  if (jr >= 0) {
    // row-free split: the samples holding the lower value, over feature 0's row
    const IdxT *row = rows_of(a, nd.src, nd.tree, 0) + nd.start;
    const uint32_t p0 = tk.chunk * a.Cf, p1 = min(nd.n, p0 + a.Cf);
    // the feature's bit column (N / 8 bytes) rather than the signatures
    const uint32_t *col = reinterpret_cast<const uint32_t *>(a.codes + fc_off(a.fcode[ax.best_f]));
    for (uint32_t pb = p0; pb < p1; pb += kFillU * kWave) {
      uint32_t e[kFillU], low[kFillU];
#pragma unroll
      for (int u = 0; u < kFillU; ++u) {
        const uint32_t p = pb + (uint32_t)u * kWave + lane;
        e[u] = (p < p1) ? (uint32_t)row[p] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int u = 0; u < kFillU; ++u) {
        const uint32_t s = sample_of((IdxT)e[u]);
        low[u] = (e[u] != 0xFFFFFFFFu) ? !((col[s >> 5] >> (s & 31u)) & 1u) : 0u;
      }
      // the lower value goes left: the smaller side's samples are marked
      bool take[kFillU];
#pragma unroll
      for (int u = 0; u < kFillU; ++u) {
        const uint32_t s = sample_of((IdxT)e[u]);
        take[u] = e[u] != 0xFFFFFFFFu && (low[u] != 0u) == small_left;
        if (take[u]) atomicOr(&bits[s >> 5], 1u << (s & 31u));
      }
      if (count) rl_add(a, e, take, cnt);
    }
  } else {
    // split feature's row: positions below nL go left, so the smaller side
    // is the range [0, nL) or [nL, n)
    const IdxT *row = rows_of(a, nd.src, nd.tree, (uint32_t)ax.best_f) + nd.start;
    const uint32_t r0 = small_left ? 0u : ax.nL, r1 = small_left ? ax.nL : nd.n;
    const uint32_t p0 = r0 + tk.chunk * a.Cf, p1 = min(r1, p0 + a.Cf);
    for (uint32_t pb = p0; pb < p1; pb += kFillU * kWave) {
      uint32_t e[kFillU];
      bool take[kFillU];
#pragma unroll
      for (int u = 0; u < kFillU; ++u) {
        const uint32_t p = pb + (uint32_t)u * kWave + lane;
        e[u] = (p < p1) ? (uint32_t)row[p] : 0xFFFFFFFFu;
        take[u] = p < p1;
      }
#pragma unroll
      for (int u = 0; u < kFillU; ++u) {
        const uint32_t s = sample_of((IdxT)e[u]);
        if (take[u]) atomicOr(&bits[s >> 5], 1u << (s & 31u));
      }
      if (count) rl_add(a, e, take, cnt);
    }
  }
}

// Fill workgroups of the tree whose frontier is [lo, hi): one per
// kFillTasksPerGroup of its fill slots [fb, fe), at most J.
#ifndef CTA_FILL_TPG
#define CTA_FILL_TPG 32
#endif
constexpr uint32_t kFillTasksPerGroup = CTA_FILL_TPG;
template <class IdxT, class LabT>
__device__ __forceinline__ uint32_t fill_groups(const FitArgs<IdxT, LabT> &a, uint32_t lo, uint32_t hi, uint32_t J,
                                                uint32_t &fb, uint32_t &fe) {
  fb = a.aux[lo].fill_base;
  const NAux last = a.aux[hi - 1];
  fe = last.fill_base + last.nch_f;
  return min(J, max(1u, ceil_div(fe - fb, kFillTasksPerGroup)));
}

template <class IdxT, class LabT>
__global__ __launch_bounds__(kFillThreads) void k_fill_lds(FitArgs<IdxT, LabT> a, uint32_t J) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  extern __shared__ uint32_t lbits[];
  const uint32_t t = blockIdx.x / J, j = blockIdx.x % J;
  const uint32_t lo = a.tlo[t], hi = a.thi[t];
  if (hi <= lo) return;  // no frontier node in this tree
  uint32_t fb, fe;
  const uint32_t Je = fill_groups(a, lo, hi, J, fb, fe);
  if (j >= Je) return;
  const uint32_t W = a.W, tid = threadIdx.x, wave = tid / kWave;
  for (uint32_t w = tid; w < W; w += kFillThreads) lbits[w] = 0;
  __syncthreads();
  // the waves' row-free counters follow the bitmap
  uint32_t *cnt = lbits + W + wave * a.Frl * (uint32_t)a.K;
  rl_zero(cnt, a.Frl * (uint32_t)a.K);
  uint32_t cur = kLeaf;
  bool cur_whole = false;
  for (uint32_t k = fb + j * kFillWaves + wave; k < fe; k += Je * kFillWaves) {
    const Task tk = a.fill_tasks[k];
    if (tk.node == kLeaf || tk.chunk >= tk.nch) continue;
    fill_task_bits(a, tk, lbits, cnt, cur, cur_whole);
  }
  if (cur != kLeaf) rl_flush(a, cur, cnt, cur_whole);
  __syncthreads();
  uint32_t *out = a.bits + (size_t)t * W;
  if (Je == 1) {
    for (uint32_t w = tid; w < W; w += kFillThreads) out[w] = lbits[w];
  } else {
    for (uint32_t w = tid; w < W; w += kFillThreads)
      if (lbits[w]) atomicOr(&out[w], lbits[w]);
  }
}

// Bitmaps too large for LDS: global atomics into the bitmaps k_select
// cleared (LDS: the waves' row-free counters).
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_fill_global(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  extern __shared__ uint32_t smem[];
  uint32_t *cnt = smem + (threadIdx.x / kWave) * a.Frl * (uint32_t)a.K;
  rl_zero(cnt, a.Frl * (uint32_t)a.K);
  uint32_t cur = kLeaf;
  bool cur_whole = false;
  WAVE_LOOP(a.ctl->n_fill) {
    const Task tk = a.fill_tasks[item];
    if (tk.node == kLeaf || tk.chunk >= tk.nch) continue;
    fill_task_bits(a, tk, a.bits + (size_t)a.front[tk.node].tree * a.W, cnt, cur, cur_whole);
  }
  if (cur != kLeaf) rl_flush(a, cur, cnt, cur_whole);
}




// Reshuffle task descriptors of node i at its scanned offsets.
template <class IdxT, class LabT>
__device__ void emit_resh_tasks(const FitArgs<IdxT, LabT> &a, uint32_t i, const NAux &ax, uint32_t bbase,
                                uint32_t sbase) {
  const uint32_t lane = lane_id();
  const uint32_t F = a.Fr;  // row slots
  if (ax.resh_ntask == 0) return;
  if (ax.resh_big) {
    for (uint32_t f = lane; f < F; f += kWave) a.big_tasks[bbase + f] = Task{i, f | (1u << 16), 0u, 1u};
  } else {
    for (uint32_t t = lane; t < ax.resh_ntask; t += kWave) {
      const uint32_t first = t * ax.rpt;
      a.small_tasks[sbase + t] = Task{i, first | (min(ax.rpt, F - first) << 16), 0u, 1u};
    }
  }
}

// True if node i's scanned ranges fit the lists.
template <class IdxT, class LabT>
__device__ __forceinline__ bool resh_fits(const FitArgs<IdxT, LabT> &a, const NAux &ax, const uint32_t *o) {
  if (ax.resh_ntask == 0) return true;
  return ax.resh_big ? (uint64_t)o[2] + ax.resh_ntask <= a.cap_big : (uint64_t)o[3] + ax.resh_ntask <= a.cap_small;
}

// Child ids, arena entries of both children, the next frontier (in parent
// order, left child first) and the reshuffle tasks of every split node.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_children_emit(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  const uint32_t id_base = a.ctl->n_nodes, sub_base = a.ctl->n_sub;
  WAVE_LOOP(a.ctl->q) {
    const uint32_t i = item;
    const NAux ax = a.aux[i];
    if (!(ax.flags & kSplit)) continue;
    const FNode nd = a.front[i];
    const uint32_t *o = a.offs + (size_t)i * kChFields;
    const uint32_t id0 = id_base + o[0];
    uint32_t slot = o[1], qslot = sub_base + o[5];
    const uint32_t nsub = ((ax.flags & kLiveL) && ax.nL <= a.sub_max ? 1u : 0u) +
                          ((ax.flags & kLiveR) && nd.n - ax.nL <= a.sub_max ? 1u : 0u);
    const uint32_t nlive = ((ax.flags & kLiveL) ? 1u : 0u) + ((ax.flags & kLiveR) ? 1u : 0u) - nsub;
    if ((uint64_t)id0 + 2 > a.cap_nodes || (uint64_t)slot + nlive > a.cap_q ||
        (uint64_t)qslot + nsub > a.cap_sub || !resh_fits(a, ax, o)) {
      if (lane == 0) atomicOr(&a.ctl->err, kErrCapacity);
      continue;
    }
    if (lane == 0) {
      a.arena.left[nd.nid] = id0;
      a.arena.right[nd.nid] = id0 + 1;
    }
    uint32_t lab[2] = {ax.lab[0], ax.lab[1]};
    if (lab[0] == kFirstSample || lab[1] == kFirstSample) {
      // leaves of fewer than min_samples_split samples: labelled by the
      // first sample of each child in feature 0's row, the first marked and
      // first unmarked sample of the parent's segment (random_tree.py:628-634)
      const IdxT *row0 = rows_of(a, nd.src, nd.tree, 0) + nd.start;
      const uint32_t *bits = a.bits + (size_t)nd.tree * a.W;
      const uint32_t inv = (ax.flags & kInvert) ? 1u : 0u;  // the marks are on the right
      bool needl = lab[0] == kFirstSample, needr = lab[1] == kFirstSample;
      for (uint32_t pb = 0; pb < nd.n && (needl || needr); pb += kWave) {
        const uint32_t p = pb + lane;
        uint32_t e = 0, side = 2;
        if (p < nd.n) {
          e = (uint32_t)row0[p];
          const uint32_t s = sample_of((IdxT)e);
          side = ((bits[s >> 5] >> (s & 31u)) & 1u) ^ inv;
        }
        const uint64_t ml = ballot(side == 1), mr = ballot(side == 0);
        if (needl && ml) {
          lab[0] = a.label((IdxT)readlane(e, (uint32_t)__ffsll((unsigned long long)ml) - 1u));
          needl = false;
        }
        if (needr && mr) {
          lab[1] = a.label((IdxT)readlane(e, (uint32_t)__ffsll((unsigned long long)mr) - 1u));
          needr = false;
        }
      }
    }
    const uint32_t cd = nd.depth + 1;
    for (uint32_t c = 0; c < 2; ++c) {
      const uint32_t id = id0 + c;
      const uint32_t cn = c ? nd.n - ax.nL : ax.nL;
      if (lane == 0) {
        a.arena.nsamp[id] = cn;
        a.arena.depth[id] = cd;
      }
      if (lab[c] != kLeaf) {
        if (lane == 0) {
          a.arena.feat[id] = -1;
          a.arena.thr[id] = 0.f;
          a.arena.left[id] = kLeaf;
          a.arena.right[id] = kLeaf;
          a.arena.value[id] = lab[c];
        }
        continue;
      }
      FNode ch;
      ch.tree = nd.tree;
      ch.start = c ? nd.start + ax.nL : nd.start;
      ch.n = cn;
      ch.nid = id;
      ch.depth = cd;
      ch.src = ax.dst;
      ch.pad0 = ch.pad1 = 0;
      if (cn <= a.sub_max) {
        // grown to its leaves by a subtree wave after the levels
        if (lane == 0) a.sub_queue[qslot] = ch;
        ++qslot;
        continue;
      }
      if (lane == 0) a.next_front[slot] = ch;
      for (uint32_t k = lane; k < K; k += kWave) {
        const uint32_t l = a.cntL[(size_t)i * K + k];
        a.next_cnt[(size_t)slot * K + k] = c ? a.cnt[(size_t)i * K + k] - l : l;
      }
      if (a.Frl) {
        // row-free counts: the smaller child's as the fill tasks counted
        // them, the other's as the parent's minus those
        const uint32_t FK = a.Frl * K;
        const bool small = (c == 0) == (ax.nL <= nd.n - ax.nL);
        const uint32_t *par = a.rlc + (size_t)i * FK, *sm = a.rl_small + (size_t)i * FK;
        uint32_t *dst = a.rlc_next + (size_t)slot * FK;
        for (uint32_t w = lane; w < FK; w += kWave) dst[w] = small ? sm[w] : par[w] - sm[w];
      }
      ++slot;
    }
    emit_resh_tasks(a, i, ax, o[2], o[3]);
  }
}

// ------------------------------------------------------------------------
// Row-free counts: per node, per row-free feature j and class k, the
// samples of class k holding the higher value of j.  The fill tasks count
// each split node's smaller side and k_children_emit derives the children's
// counts; the roots (and frontiers set by the host) are counted here, a
// wave per chunk of a node's feature-0 row into LDS (one atomic per set
// signature bit), then added to the node's counts in global memory.
// ------------------------------------------------------------------------
constexpr uint32_t kRlCounters = 1024;  // 2 x Frl x K at most (per-wave LDS counters)

// Row-free counts of every sample at positions [p0, p1) of node nd's
// feature-0 row, into the wave's counters.
template <class IdxT, class LabT>
__device__ __forceinline__ void rl_count_chunk(const FitArgs<IdxT, LabT> &a, const FNode &nd, uint32_t p0, uint32_t p1,
                                               uint32_t *cnt) {
  const uint32_t lane = lane_id();
  const IdxT *row = rows_of(a, nd.src, nd.tree, 0) + nd.start;
  constexpr int U = 4;  // positions per lane in flight
  for (uint32_t pb = p0; pb < p1; pb += U * kWave) {
    uint32_t e[U];
    bool take[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t p = pb + (uint32_t)u * kWave + lane;
      take[u] = p < p1;
      e[u] = take[u] ? (uint32_t)row[p] : 0u;
    }
    rl_add(a, e, take, cnt);
  }
}

// The current frontier's counts from scratch (roots, frontiers set by the
// host): wave w of block (x, i) counts chunks x kWavesPerBlock + w, then
// every gridDim.x kWavesPerBlock-th, of node i and adds them to rlc, which
// must be zero.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_rl_front(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t FK = a.Frl * (uint32_t)a.K;
  const uint32_t q = a.ctl->q;
  uint32_t *cnt = smem + wave * kRlCounters;
  for (uint32_t i = blockIdx.y; i < q; i += gridDim.y) {
    const FNode nd = a.front[i];
    const uint32_t nchunk = ceil_div(nd.n, a.Cf);
    const uint32_t c0 = blockIdx.x * kWavesPerBlock + wave;
    if (c0 >= nchunk) continue;
    // every chunk of this wave counted, then one flush (few atomics on each
    // of the node's counters)
    for (uint32_t w = lane; w < FK; w += kWave) cnt[w] = 0;
    wave_sync();
    for (uint32_t c = c0; c < nchunk; c += gridDim.x * kWavesPerBlock)
      rl_count_chunk(a, nd, c * a.Cf, min(nd.n, (c + 1) * a.Cf), cnt);
    wave_sync();
    uint32_t *dst = const_cast<uint32_t *>(a.rlc) + (size_t)i * FK;
    for (uint32_t w = lane; w < FK; w += kWave)
      if (cnt[w]) atomicAdd(&dst[w], cnt[w]);
    wave_sync();
  }
}

// Reshuffle tasks only (bootstrap compaction, reshuffle hook).
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_emit_resh(FitArgs<IdxT, LabT> a) {
  if (a.ctl->err) return;  // a list overflowed: the fit fails, nothing more is read
  WAVE_LOOP(a.ctl->q) {
    const NAux ax = a.aux[item];
    if (!(ax.flags & kSplit)) continue;
    const uint32_t *o = a.offs + (size_t)item * kChFields;
    if (!resh_fits(a, ax, o)) {
      if (lane_id() == 0) atomicOr(&a.ctl->err, kErrCapacity);
      continue;
    }
    emit_resh_tasks(a, item, ax, o[2], o[3]);
  }
}

// ------------------------------------------------------------------------
// Reshuffle: stable partition of every row of the split nodes.
//
// k_plan reads each tree's task ranges off the children-scan offsets at its
// frontier range and gives the tree workgroups in proportion to its
// elements, so that the level's work spreads over about `target`
// workgroups (each moving between kWgElemsMin and kWgElemsMax elements).
// It also clears the tree ranges for the next level.  k_resh runs
// 1,024-thread workgroups, each bound to one tree: it stages the tree's side
// bitmap in LDS (when it fits), then moves its share of the tree's big tasks
// cooperatively (kReshTile-element tiles, one barrier per tile, the left
// count carried across tiles) and its share of the small tasks one wave per
// task.  Loads and stores are striped (lane-consecutive), and the stable
// ranks come from one ballot per 64 elements.
// ------------------------------------------------------------------------
constexpr int kPlanThreads = 1024;

// End of a level: the children-scan totals become the next frontier size
// and the ids in use (also recorded per level in lv_q / lv_nodes); the
// spread byte counters are folded into Ctl.  Run by the first kByteKinds
// waves of one workgroup, one wave per counter kind.
__device__ __forceinline__ void ctl_next_body(Ctl *ctl, unsigned long long *bslot, uint32_t tid) {
  if (tid >= kByteKinds * kWave) return;
  const uint32_t kind = tid / kWave, slot = tid % kWave;
  unsigned long long v = 0;
  for (uint32_t s = slot; s < kByteSlots; s += kWave) {
    unsigned long long *p = bslot + ((size_t)kind * kByteSlots + s) * kByteStride;
    v += *p;
    *p = 0;
  }
  v = wave_sum64(v);
  if (slot == 0 && v) (&ctl->bytes_scan)[kind] += v;
  if (tid == 0) {
    ctl->q = ctl->c_live;
    ctl->n_nodes += ctl->c_ids;
    ctl->n_sub += ctl->c_sub;
    ctl->c_live = 0;
    ctl->c_ids = 0;
    ctl->c_sub = 0;
    // per-level record, so the host can read several levels at once
    const uint32_t d = (ctl->depth + 1) % kLevelRing;
    ctl->depth += 1;
    ctl->lv_nodes[d] = ctl->n_nodes;
    ctl->lv_q[d] = ctl->q;
  }
}

// With end_level set, the workgroup then ends the level (ctl_next_body):
// the reshuffle launches that follow read only the plan.  With a.ctl_host
// set it then copies the control block there.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kPlanThreads) void k_plan(FitArgs<IdxT, LabT> a, uint32_t target, uint32_t end_level) {
  constexpr int kW = kPlanThreads / kWave;
  __shared__ uint32_t ws[2][kW];
  __shared__ uint32_t carry[2];
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const uint32_t q = a.ctl->q, T = a.T;
  const uint32_t tot_big = a.ctl->c_big, tot_small = a.ctl->c_small, tot_tiles = a.ctl->c_tiles;
  const unsigned long long total =
      (unsigned long long)tot_tiles * kReshTile + (unsigned long long)tot_small * a.Cr;
  const unsigned long long per =
      min(kWgElemsMax, max(kWgElemsMin, total / max(1u, target)));
  if (tid < 2) carry[tid] = 0;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < T; t0 += kPlanThreads) {
    const uint32_t t = t0 + tid;
    uint32_t wgs[2] = {0u, 0u};
    if (t < T) {
      const uint32_t lo = a.tlo[t], hi = a.thi[t];
      uint32_t bb = 0, be = 0, sb = 0, se = 0, tb = 0, te = 0;
      if (hi > lo) {
        const uint32_t *ol = a.offs + (size_t)lo * kChFields;
        bb = ol[2];
        sb = ol[3];
        tb = ol[4];
        if (hi < q) {
          const uint32_t *oh = a.offs + (size_t)hi * kChFields;
          be = oh[2];
          se = oh[3];
          te = oh[4];
        } else {
          be = tot_big;
          se = tot_small;
          te = tot_tiles;
        }
      }
      uint32_t *pl = a.plan + (size_t)t * 4;
      pl[0] = bb;
      pl[1] = be;
      pl[2] = sb;
      pl[3] = se;
      const unsigned long long eb = (unsigned long long)(te - tb) * kReshTile;
      const unsigned long long es = (unsigned long long)(se - sb) * a.Cr;
      if (eb) wgs[0] = (uint32_t)min(4096ull, (eb + per - 1) / per);
      if (es) wgs[1] = (uint32_t)min(4096ull, (es + per - 1) / per);
      a.tlo[t] = 0;
      a.thi[t] = 0;
    }
    uint32_t inc[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      inc[k] = wave_incl_scan(wgs[k]);
      if (lane == kWave - 1) ws[k][wave] = inc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      uint32_t off = carry[k] + inc[k] - wgs[k];
      for (uint32_t w = 0; w < wave; ++w) off += ws[k][w];
      if (t < T) a.wg_off[(size_t)k * (T + 1) + t] = off;
    }
    __syncthreads();
    if (tid < 2) {
      uint32_t sum = 0;
      for (int w = 0; w < kW; ++w) sum += ws[tid][w];
      carry[tid] += sum;
    }
    __syncthreads();
  }
  if (tid < 2) {
    uint32_t P = carry[tid];
    if (tot_big > a.cap_big || tot_small > a.cap_small || a.ctl->err) P = 0;
    a.wg_off[(size_t)tid * (T + 1) + T] = P;
    a.ctl->n_plan[tid] = P;
  }
  if (end_level) {
    __syncthreads();  // every thread has read this level's counts
    ctl_next_body(a.ctl, a.bslot, tid);
    if (a.ctl_host) {
      // the control block for the host, which reads it after an event on
      // this launch while the level's reshuffle still runs
      __syncthreads();
      const uint32_t *src = reinterpret_cast<const uint32_t *>(a.ctl);
      uint32_t *dst = reinterpret_cast<uint32_t *>(a.ctl_host);
      for (uint32_t w = tid; w < (uint32_t)(sizeof(Ctl) / 4); w += kPlanThreads) dst[w] = src[w];
      __threadfence_system();
    }
  }
}

// Global-memory pointers held in scalar registers, with row elements
// addressed by 32-bit byte offsets (rows hold at most 2^22 elements), so
// loads and stores take the scalar-base global form and keep vector
// registers free.
template <class T>
using gptr = __attribute__((address_space(1))) T *;

template <class T>
__device__ __forceinline__ gptr<T> sgpr_ptr(T *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  return reinterpret_cast<gptr<T>>(((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v));
}
template <class T>
__device__ __forceinline__ T ld32(gptr<const T> base, uint32_t i) {
  return *reinterpret_cast<gptr<const T>>(reinterpret_cast<gptr<const char>>(base) + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(gptr<T> base, uint32_t i, T v) {
  *reinterpret_cast<gptr<T>>(reinterpret_cast<gptr<char>>(base) + i * (uint32_t)sizeof(T)) = v;
}

// One big task: all rows' segment [0, n) of one row, moved by the workgroup.
// `par` selects the half of wcnt a tile writes; it alternates across tiles
// and across consecutive tasks, so a tile never overwrites counts that a
// slower wave of the previous tile may still be reading.  Tiles are
// software-pipelined over two register sets: the next tile's loads are
// issued before this tile's stores, so neither waits for the other.
template <class IdxT, class LabT, class Side>
__device__ void resh_big(const FitArgs<IdxT, LabT> &a, const Task &tk, const Side &side, uint32_t *wcnt,
                         uint32_t &par) {
  const uint32_t node = uniform(tk.node);
  const FNode nd = a.front[node];
  const NAux ax = a.aux[node];
  const uint32_t f = uniform(tk.first_count & 0xFFFFu);
  const gptr<const IdxT> in = sgpr_ptr(rows_of_slot(a, nd.src, nd.tree, f) + nd.start);
  const gptr<IdxT> out = sgpr_ptr(rows_out(a, ax.dst, nd.tree, f) + nd.start);
  const bool liveL = uniform(ax.flags) & kLiveL, liveR = uniform(ax.flags) & kLiveR;
  const uint32_t inv = (uniform(ax.flags) & kInvert) ? 1u : 0u;  // marks on the right child
  const uint32_t n = uniform(nd.n), nL = uniform(ax.nL);
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t carry = 0;
  // positions past n repeat the last element and are masked
  auto load = [&](IdxT(&e)[kBigE], uint32_t base) {
#pragma unroll
    for (int k = 0; k < kBigE; ++k) e[k] = ld32(in, min(base + (uint32_t)k * kReshThreads + tid, n - 1));
  };
  // moves the tile at `base` held in e; with two register sets, loads the
  // next tile into nx before this tile's stores, else after them into e
  auto step = [&](IdxT(&e)[kBigE], IdxT(&nx)[kBigE], uint32_t base) -> bool {
    // Every e[k] is a valid entry, so the side lookups are issued together,
    // unconditionally, and masked afterwards (branch-free: a lookup under a
    // condition would wait for its own load and stall the next tile's).
    uint32_t sv[kBigE];
#pragma unroll
    for (int k = 0; k < kBigE; ++k) sv[k] = side(e[k]);
    uint32_t sdm = 0;  // bit k: element k goes left
#pragma unroll
    for (int k = 0; k < kBigE; ++k) {
      const uint32_t p = base + (uint32_t)k * kReshThreads + tid;
      sdm |= ((sv[k] ^ inv) & (p < n ? 1u : 0u)) << k;
    }
    uint64_t mk[kBigE];
#pragma unroll
    for (int k = 0; k < kBigE; ++k) {
      mk[k] = ballot((sdm >> k) & 1u);
      if (lane == 0) wcnt[par * kSlots + k * kReshWaves + wave] = (uint32_t)__popcll(mk[k]);
    }
    __syncthreads();
    // exclusive prefix over the (sub-tile, wave) counts in position order,
    // kSpl consecutive slots per lane; the slots of this wave's column
    // (k * kReshWaves + wave) sit at index wave % kSpl of their lane
    uint32_t v[kSpl], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSpl; ++j) {
      v[j] = wcnt[par * kSlots + kSpl * lane + j];
      sum += v[j];
    }
    const uint32_t inc = wave_incl_scan(sum);
    const uint32_t total = readlane(inc, 63);
    const uint32_t js = wave % kSpl;
    uint32_t esel = inc - sum;
#pragma unroll
    for (uint32_t j = 0; j + 1 < kSpl; ++j) esel += (j < js) ? v[j] : 0u;
    const bool more = base + kReshTile < n;
    if (kBigPipe && more) load(nx, base + kReshTile);
#pragma unroll
    for (int k = 0; k < kBigE; ++k) {
      const uint32_t slot = (uint32_t)k * kReshWaves + wave;
      const uint32_t before = carry + readlane(esel, slot / kSpl);
      const uint32_t p = base + (uint32_t)k * kReshThreads + tid;
      const bool left = (sdm >> k) & 1u;
      const uint32_t lb = before + mbcnt(mk[k]);
      // one store per element: left at its rank, right after the nL left ones
      if (p < n && (left ? liveL : liveR)) st32(out, left ? lb : nL + (p - lb), e[k]);
    }
    if (!kBigPipe && more) load(e, base + kReshTile);
    carry += total;
    par ^= 1u;
    return more;
  };
  if constexpr (kBigPipe) {
    IdxT A[kBigE], B[kBigE];
    load(A, 0);
    for (uint32_t base = 0;; base += 2 * kReshTile) {
      if (!step(A, B, base)) break;
      if (!step(B, A, base + kReshTile)) break;
    }
  } else {
    IdxT A[kBigE];
    load(A, 0);
    for (uint32_t base = 0; step(A, A, base); base += kReshTile) {
    }
  }
}

// Row slots [first, end) of the node of tree `tree` whose segment
// [start, start + n) sits in rows `src`, partitioned by one wave into rows
// `dst`: samples whose side(s) ^ inv is 1 to [start, start + nL), the others
// after them (flags: kLiveL / kLiveR, the children whose rows are written,
// and kInvert).  A row of n <= 64 * kReshE / 2 samples takes c = ceil(n / 64)
// pieces, so kReshE / c rows are loaded together (one batch of kReshE
// striped loads in flight) before their ranks are taken in order.
// Consecutive rows of a tree are N elements apart.
template <class IdxT, class LabT, class Side>
__device__ void resh_rows(const FitArgs<IdxT, LabT> &a, uint32_t tree, uint32_t start, uint32_t src, uint32_t dst,
                          uint32_t n, uint32_t nL, uint32_t flags, uint32_t first, uint32_t end, const Side &side) {
  const bool liveL = flags & kLiveL, liveR = flags & kLiveR;
  const uint32_t inv = (flags & kInvert) ? 1u : 0u;  // marks on the right child
  const uint32_t N = (uint32_t)a.N;
  const uint32_t lane = lane_id();
  const uint32_t c = ceil_div(n, kWave);
  if (c <= (uint32_t)kReshE / 2) {
    const uint32_t R = (uint32_t)kReshE / c;  // rows per batch
    for (uint32_t f0 = first; f0 < end; f0 += R) {
      const gptr<IdxT> out = sgpr_ptr(rows_out(a, dst, tree, f0) + start);
      const uint32_t rows = min(R, end - f0);
      IdxT e[kReshE];
      uint32_t r = 0, pc = 0;
#pragma unroll
      for (int k = 0; k < kReshE; ++k) {
        const uint32_t p = pc * kWave + lane;
        // source rows by slot (the presorted rows are one per feature)
        const gptr<const IdxT> in = sgpr_ptr(rows_of_slot(a, src, tree, f0 + min(r, rows - 1)) + start);
        e[k] = ld32(in, min(p, n - 1));  // out-of-range pieces are masked
        if (++pc == c) {
          pc = 0;
          ++r;
        }
      }
      // every e[k] is a valid entry: lookups issued together, masked after
      uint32_t sv[kReshE];
#pragma unroll
      for (int k = 0; k < kReshE; ++k) sv[k] = side(e[k]);
      uint32_t sdm = 0;
      r = 0;
      pc = 0;
#pragma unroll
      for (int k = 0; k < kReshE; ++k) {
        const uint32_t p = pc * kWave + lane;
        sdm |= ((sv[k] ^ inv) & (r < rows && p < n ? 1u : 0u)) << k;
        if (++pc == c) {
          pc = 0;
          ++r;
        }
      }
      uint32_t carry = 0;
      r = 0;
      pc = 0;
#pragma unroll
      for (int k = 0; k < kReshE; ++k) {
        const uint32_t p = pc * kWave + lane;
        const bool left = (sdm >> k) & 1u;
        const uint64_t m = ballot(left);
        const uint32_t lb = carry + mbcnt(m);
        if (r < rows && p < n && (left ? liveL : liveR)) st32(out, r * N + (left ? lb : nL + (p - lb)), e[k]);
        carry += (uint32_t)__popcll(m);
        if (++pc == c) {
          pc = 0;
          ++r;
          carry = 0;
        }
      }
    }
    return;
  }
  // Pieces of kWave * kReshE positions, row by row; the next piece's loads
  // are issued before this piece's stores (two register sets).
  constexpr uint32_t kPiece = kWave * kReshE;
  if (first >= end) return;
  const size_t rstride = (size_t)a.N;  // consecutive row slots of a tree are N entries apart
  const gptr<IdxT> out0 = sgpr_ptr(rows_out(a, dst, tree, first) + start);
  auto load = [&](IdxT(&e)[kReshE], uint32_t r, uint32_t base) {
    const gptr<const IdxT> in = sgpr_ptr(rows_of_slot(a, src, tree, first + r) + start);
#pragma unroll
    for (int k = 0; k < kReshE; ++k) e[k] = ld32(in, min(base + (uint32_t)k * kWave + lane, n - 1));
  };
  uint32_t carry = 0;
  // moves the piece (row r, base) held in e; loads the next piece into nx first
  auto step = [&](IdxT(&e)[kReshE], IdxT(&nx)[kReshE], uint32_t &r, uint32_t &base) -> bool {
    uint32_t r2 = r, b2 = base + kPiece;
    if (b2 >= n) {
      b2 = 0;
      ++r2;
    }
    const bool more = first + r2 < end;
    if (more) load(nx, r2, b2);
    const gptr<IdxT> out = out0 + r * rstride;
    uint32_t sv[kReshE];
#pragma unroll
    for (int k = 0; k < kReshE; ++k) sv[k] = side(e[k]);
    uint32_t sdm = 0;
#pragma unroll
    for (int k = 0; k < kReshE; ++k) {
      const uint32_t p = base + (uint32_t)k * kWave + lane;
      sdm |= ((sv[k] ^ inv) & (p < n ? 1u : 0u)) << k;
    }
#pragma unroll
    for (int k = 0; k < kReshE; ++k) {
      const uint32_t p = base + (uint32_t)k * kWave + lane;
      const bool left = (sdm >> k) & 1u;
      const uint64_t m = ballot(left);
      const uint32_t lb = carry + mbcnt(m);
      if (p < n && (left ? liveL : liveR)) st32(out, left ? lb : nL + (p - lb), e[k]);
      carry += (uint32_t)__popcll(m);
    }
    if (r2 != r) carry = 0;
    r = r2;
    base = b2;
    return more;
  };
  IdxT A[kReshE], B[kReshE];
  uint32_t r = 0, base = 0;
  load(A, 0, 0);
  for (;;) {
    if (!step(A, B, r, base)) break;
    if (!step(B, A, r, base)) break;
  }
}

// One small task: rows [first, first+count) of a node, moved by one wave.
template <class IdxT, class LabT, class Side>
__device__ void resh_small(const FitArgs<IdxT, LabT> &a, const Task &tk, const Side &side) {
  const uint32_t node = uniform(tk.node);
  const FNode nd = a.front[node];
  const NAux ax = a.aux[node];
  const uint32_t first = uniform(tk.first_count & 0xFFFFu), count = uniform(tk.first_count >> 16);
  resh_rows(a, uniform(nd.tree), uniform(nd.start), uniform(nd.src), uniform(ax.dst), uniform(nd.n), uniform(ax.nL),
            uniform(ax.flags), first, first + count, side);
}

// One launch moves both plans: plan entries [0, P_big) are big-task work,
// the next P_small small-task work, so the small tasks fill the CUs that the
// big ones leave while a level's last big tasks finish.  kLds: the tree's
// bitmap is staged in LDS (N/8 bytes); otherwise it is read from global
// memory.
template <class IdxT, class LabT, bool kLds>
__global__ __launch_bounds__(kReshThreads, 8) void k_resh(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t lds[];
  uint32_t *wcnt = lds + (kLds ? ((a.W + 3u) & ~3u) : 0u);  // 2 x kSlots counts
  const uint32_t Pb = a.ctl->n_plan[0], P = Pb + a.ctl->n_plan[1];
  const uint32_t T = a.T;
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t loaded = 0xFFFFFFFFu, par = 0;
  for (uint32_t e = blockIdx.x; e < P; e += gridDim.x) {
    const bool big = e < Pb;
    const uint32_t pe = big ? e : e - Pb;
    // tree of this plan entry: last t with wg_off[t] <= pe
    const uint32_t *wg_off = a.wg_off + (big ? 0u : T + 1);
    uint32_t lo = 0, hi = T;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (wg_off[mid] <= pe) lo = mid;
      else hi = mid;
    }
    const uint32_t t = lo;
    const uint32_t j = pe - wg_off[t], wgs = wg_off[t + 1] - wg_off[t];
    const uint32_t *gbits = a.bits + (size_t)t * a.W;
    if (kLds && t != loaded) {
      __syncthreads();
      for (uint32_t w = threadIdx.x; w < a.W; w += kReshThreads) lds[w] = gbits[w];
      __syncthreads();
      loaded = t;
    }
    auto side = [&](IdxT s) -> uint32_t {
      const uint32_t u = sample_of(s);
      if constexpr (kLds) return (lds[u >> 5] >> (u & 31u)) & 1u;
      else return (gbits[u >> 5] >> (u & 31u)) & 1u;
    };
    const uint32_t *pl = a.plan + (size_t)t * 4;
    if (big) {
      const uint32_t b0 = pl[0], nb = pl[1] - pl[0];
      for (uint32_t bt = j; bt < nb; bt += wgs) resh_big(a, a.big_tasks[b0 + bt], side, wcnt, par);
    } else {
      const uint32_t s0 = pl[2], ns = pl[3] - pl[2];
      for (uint32_t st = j * kReshWaves + wave; st < ns; st += wgs * kReshWaves)
        resh_small(a, a.small_tasks[s0 + st], side);
    }
  }
}

// ------------------------------------------------------------------------
// Bootstrap and roots (random_forest.py:113-144, bootstrap_fill.cu,
// bootstrap_reshuffle.cu).
// ------------------------------------------------------------------------
__global__ void k_fill_u8(uint8_t *p, size_t n, uint8_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Sets the bit of every in-bag id of every batch tree (bitmaps zeroed first).
__global__ void k_boot_mark(const uint32_t *ids, const uint64_t *off, uint32_t T, uint32_t W, uint32_t *bits) {
  const uint32_t t = blockIdx.y;
  if (t >= T) return;
  const uint64_t a = off[t], b = off[t + 1];
  for (uint64_t j = a + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < b; j += (uint64_t)gridDim.x * blockDim.x)
    atomicOr(&bits[(size_t)t * W + (ids[j] >> 5)], 1u << (ids[j] & 31u));
}

// In-bag bitmaps from the canonical generator: bit s of tree t is set when
// some draw of tree (tree_base + t) lands on s (bitmaps zeroed first).
__global__ void k_boot_gen(uint64_t seed, uint32_t tree_base, uint32_t N, uint32_t W, uint32_t *bits) {
  const uint32_t t = blockIdx.y;
  uint32_t *b = bits + (size_t)t * W;
  const uint32_t nblk = (N + 3) / 4;
  for (uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x; blk < nblk; blk += gridDim.x * blockDim.x) {
    uint32_t w[4];
    boot_block(seed, tree_base + t, blk, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (blk * 4 + (uint32_t)k >= N) break;
      const uint32_t s = draw_to_sample(w[k], N);
      atomicOr(&b[s >> 5], 1u << (s & 31u));
    }
  }
}

// The same bitmaps built in LDS, one 1,024-thread workgroup per tree: every
// word is written once, so no clearing and no global atomics (N / 8 bytes
// must fit the workgroup's LDS).
__global__ __launch_bounds__(1024) void k_boot_gen_lds(uint64_t seed, uint32_t tree_base, uint32_t N, uint32_t W,
                                                       uint32_t *bits) {
  extern __shared__ uint32_t lb[];
  const uint32_t t = blockIdx.x;
  for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) lb[w] = 0;
  __syncthreads();
  const uint32_t nblk = (N + 3) / 4;
  for (uint32_t blk = threadIdx.x; blk < nblk; blk += blockDim.x) {
    uint32_t w[4];
    boot_block(seed, tree_base + t, blk, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (blk * 4 + (uint32_t)k >= N) break;
      const uint32_t s = draw_to_sample(w[k], N);
      atomicOr(&lb[s >> 5], 1u << (s & 31u));
    }
  }
  __syncthreads();
  uint32_t *b = bits + (size_t)t * W;
  for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) b[w] = lb[w];
}

// Set bits of every tree's bitmap (the in-bag counts).
__global__ void k_bits_count(const uint32_t *bits, uint32_t W, uint32_t *counts) {
  const uint32_t t = blockIdx.y;
  const uint32_t *b = bits + (size_t)t * W;
  uint32_t c = 0;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W; w += gridDim.x * blockDim.x) c += __popc(b[w]);
  c = wave_sum32(c);
  if (lane_id() == 0 && c) atomicAdd(&counts[t], c);
}

// Class counts of every tree's root from its in-bag bitmap.
template <class LabT>
__global__ void k_root_counts_bits(const LabT *y, const uint32_t *bits, uint32_t N, uint32_t W, uint32_t K,
                                   uint32_t *cnt) {
  extern __shared__ uint32_t h[];
  const uint32_t t = blockIdx.y;
  const uint32_t *b = bits + (size_t)t * W;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    if ((b[i >> 5] >> (i & 31u)) & 1u) atomicAdd(&h[y[i]], 1u);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (h[k]) atomicAdd(&cnt[(size_t)t * K + k], h[k]);
}

// Class counts of every batch tree's root: over its ids, or over all N.
template <class LabT>
__global__ void k_root_counts(const LabT *y, const uint32_t *ids, const uint64_t *off, uint32_t T, uint32_t N,
                              uint32_t K, uint32_t *cnt) {
  extern __shared__ uint32_t h[];
  const uint32_t t = blockIdx.y;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
  __syncthreads();
  const uint64_t a = ids ? off[t] : 0, b = ids ? off[t + 1] : N;
  for (uint64_t j = a + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < b; j += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[y[ids ? ids[j] : (uint32_t)j]], 1u);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (h[k]) atomicAdd(&cnt[(size_t)t * K + k], h[k]);
}

// Pseudo-frontier of the bootstrap compaction: tree t's presorted rows
// (src 0, N samples) keep the in-bag samples (bit set) in buffer 0.  One
// node per tree, so the frontier range of tree t is [t, t + 1).
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_boot_tasks(FitArgs<IdxT, LabT> a, const uint64_t *off, uint32_t T,
                                                       FNode *front) {
  const uint32_t lane = lane_id();
  unsigned long long bytes = 0;
  WAVE_LOOP(T) {
    const uint32_t t = item;
    if (lane == 0) {
      FNode nd;
      nd.tree = t;
      nd.start = 0;
      nd.n = (uint32_t)a.N;
      nd.nid = 0;
      nd.depth = 0;
      nd.src = 0;
      nd.pad0 = nd.pad1 = 0;
      front[t] = nd;
      NAux ax = {};
      ax.best_f = -1;
      ax.nL = (uint32_t)(off[t + 1] - off[t]);
      ax.flags = kSplit | kLiveL;
      ax.dst = 1;
      uint32_t v[kChFields] = {0u, 0u, 0u, 0u, 0u};
      resh_shape((uint32_t)a.N, a.Fr, a.Cr, ax, v);
      uint32_t *o = a.vals + (size_t)t * kChFields;
#pragma unroll
      for (int f = 0; f < kChFields; ++f) o[f] = v[f];
      a.aux[t] = ax;
      a.tlo[t] = t;
      a.thi[t] = t + 1;
      bytes += (unsigned long long)a.Fr * a.N * (sizeof(IdxT) + 1) + (unsigned long long)a.Fr * ax.nL * sizeof(IdxT);
    }
  }
  flush_bytes(a, kBytesBoot, bytes);
}

// Roots: tree t's root (arena id t) is a leaf if pure, else enters the
// frontier; one workgroup, so the frontier is in tree order.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kPlanThreads) void k_roots(FitArgs<IdxT, LabT> a, const uint32_t *root_cnt,
                                                        const uint64_t *off, uint32_t T, uint32_t src) {
  constexpr int kW = kPlanThreads / kWave;
  __shared__ uint32_t ws[kW];
  __shared__ uint32_t carry;
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const uint32_t K = (uint32_t)a.K;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < T; t0 += kPlanThreads) {
    const uint32_t t = t0 + tid;
    uint32_t live = 0, n = 0;
    if (t < T) {
      n = off ? (uint32_t)(off[t + 1] - off[t]) : (uint32_t)a.N;
      uint32_t nz = 0, only = 0;
      for (uint32_t k = 0; k < K; ++k)
        if (root_cnt[(size_t)t * K + k]) {
          ++nz;
          only = k;
        }
      a.arena.nsamp[t] = n;
      a.arena.depth[t] = 0;
      // The reference never tests the root for purity (random_tree.py:482-486);
      // a pure root is a leaf here.  min_samples_split applies as at any node;
      // a positive max_depth never caps the root.
      uint32_t leaf_label = kLeaf;
      if (nz <= 1) leaf_label = only;
      else if ((int64_t)n < (int64_t)a.mss) leaf_label = a.label(rows_of(a, src, t, 0)[0]);
      if (leaf_label != kLeaf) {
        a.arena.feat[t] = -1;
        a.arena.thr[t] = 0.f;
        a.arena.left[t] = kLeaf;
        a.arena.right[t] = kLeaf;
        a.arena.value[t] = leaf_label;
      } else {
        live = 1;
      }
    }
    const uint32_t inc = wave_incl_scan(live);
    if (lane == kWave - 1) ws[wave] = inc;
    __syncthreads();
    uint32_t slot = carry + inc - live;
    for (uint32_t w = 0; w < wave; ++w) slot += ws[w];
    if (live) {
      FNode nd;
      nd.tree = t;
      nd.start = 0;
      nd.n = n;
      nd.nid = t;
      nd.depth = 0;
      nd.src = src;
      nd.pad0 = nd.pad1 = 0;
      a.next_front[slot] = nd;
      for (uint32_t k = 0; k < K; ++k) a.next_cnt[(size_t)slot * K + k] = root_cnt[(size_t)t * K + k];
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t sum = 0;
      for (int w = 0; w < kW; ++w) sum += ws[w];
      carry += sum;
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.ctl->c_live = carry;
    a.ctl->c_ids = 0;
  }
}

// ------------------------------------------------------------------------
// Canonical export: pre-order numbering of every tree on the device.
// Nodes are bucketed by depth; subtree sizes are summed bottom-up one depth
// at a time, pre-order positions assigned top-down (left child = parent + 1,
// right child = parent + 1 + size of the left subtree), then one scatter
// writes the canonical arrays and the flattened predict table.
// ------------------------------------------------------------------------
__global__ void k_depth_hist(const uint32_t *depth, uint32_t n, uint32_t *cnt) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&cnt[depth[i]], 1u);
}

// In-place exclusive prefix of v[0..n) with the total in v[n] (one thread).
__global__ void k_prefix_serial(uint32_t *v, uint32_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t run = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  v[n] = run;
}

__global__ void k_depth_scatter(const uint32_t *depth, uint32_t n, uint32_t *cursor, uint32_t *order) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    order[atomicAdd(&cursor[depth[i]], 1u)] = i;
}

// order == nullptr: the ids of a depth are the contiguous range [b, e).
__global__ void k_subtree_size(const uint32_t *order, uint32_t b, uint32_t e, Arena ar, uint32_t *size) {
  for (uint32_t k = b + blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
    const uint32_t id = order ? order[k] : k;
    size[id] = ar.feat[id] < 0 ? 1u : 1u + size[ar.left[id]] + size[ar.right[id]];
  }
}

__global__ void k_root_sizes(const uint32_t *size, uint32_t T, uint32_t *tsz) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) tsz[t] = size[t];
}

// pos: position inside the tree; tr: tree of the node.  Roots are set first.
__global__ void k_preorder(const uint32_t *order, uint32_t b, uint32_t e, Arena ar, const uint32_t *size,
                           uint32_t *pos, uint32_t *tr) {
  for (uint32_t k = b + blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
    const uint32_t id = order ? order[k] : k;
    if (ar.feat[id] < 0) continue;
    const uint32_t l = ar.left[id], r = ar.right[id];
    pos[l] = pos[id] + 1;
    pos[r] = pos[id] + 1 + size[l];
    tr[l] = tr[id];
    tr[r] = tr[id];
  }
}

__global__ void k_root_pos(uint32_t T, uint32_t *pos, uint32_t *tr) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
    pos[t] = 0;
    tr[t] = t;
  }
}

// Subtree sizes and pre-order positions of the trees of a batch whose node
// ids were allocated one depth at a time: depth d holds ids
// [doff[d], doff[d + 1]), grouped by tree in tree order, and a split node's
// children are allocated as a pair in the order of their parents, so the
// ids of tree t at depth d + 1 run from the left child of its first split
// node at depth d to the right child of its last.  Workgroup t handles tree
// t (root id t): its id range at every depth (kept in LDS, D + 1 of them),
// then sizes bottom-up and positions top-down, one depth at a time, each
// depth's writes made visible to the next by the workgroup barrier.
constexpr int kLayoutThreads = 1024;
constexpr uint32_t kLayoutMaxDepth = 4095;  // deeper batches take the per-depth kernels
__global__ __launch_bounds__(kLayoutThreads) void k_tree_layout(Arena ar, uint32_t D, uint32_t *size, uint32_t *pos,
                                                                uint32_t *tr) {
  extern __shared__ uint32_t rng[];  // (lo, hi) per depth
  __shared__ uint32_t red[2][kLayoutThreads / 64];
  const uint32_t t = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  uint32_t lo = t, hi = t + 1, dmax = 0;
  for (uint32_t d = 0;; ++d) {
    if (tid == 0) {
      rng[2 * d] = lo;
      rng[2 * d + 1] = hi;
    }
    dmax = d;
    if (d == D) break;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (uint32_t id = lo + tid; id < hi; id += kLayoutThreads)
      if (ar.feat[id] >= 0) {
        mn = min(mn, ar.left[id]);
        mx = max(mx, ar.right[id]);
      }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, kWave));
      mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, kWave));
    }
    if (lane == 0) {
      red[0][w] = mn;
      red[1][w] = mx;
    }
    __syncthreads();
    mn = 0xFFFFFFFFu;
    mx = 0;
#pragma unroll
    for (int k = 0; k < kLayoutThreads / 64; ++k) {
      mn = min(mn, red[0][k]);
      mx = max(mx, red[1][k]);
    }
    __syncthreads();
    if (mn == 0xFFFFFFFFu) break;  // no split node: the tree ends at depth d
    lo = mn;
    hi = mx + 1;
  }
  __syncthreads();
  for (uint32_t d = dmax + 1; d-- > 0;) {
    const uint32_t a = rng[2 * d], b = rng[2 * d + 1];
    for (uint32_t id = a + tid; id < b; id += kLayoutThreads) {
      size[id] = ar.feat[id] < 0 ? 1u : 1u + size[ar.left[id]] + size[ar.right[id]];
      tr[id] = t;
    }
    __syncthreads();
  }
  if (tid == 0) pos[t] = 0;
  __syncthreads();
  for (uint32_t d = 0; d < dmax; ++d) {
    const uint32_t a = rng[2 * d], b = rng[2 * d + 1];
    for (uint32_t id = a + tid; id < b; id += kLayoutThreads) {
      if (ar.feat[id] < 0) continue;
      const uint32_t l = ar.left[id], r = ar.right[id], p = pos[id];
      pos[l] = p + 1;
      pos[r] = p + 1 + size[l];
    }
    __syncthreads();
  }
}

struct CanonOut {
  int32_t *feat;
  float *thr;
  uint32_t *left, *right, *value, *nsamp, *depth;
  PNode *pnode;
};

__global__ void k_canon_scatter(uint32_t n, Arena ar, const uint32_t *size, const uint32_t *pos, const uint32_t *tr,
                                const uint32_t *tree_off, CanonOut o) {
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
    const uint32_t base = tree_off[tr[id]];
    const uint32_t p = pos[id], g = base + p;
    const bool leaf = ar.feat[id] < 0;
    const uint32_t rl = leaf ? 0u : p + 1 + size[ar.left[id]];
    o.feat[g] = leaf ? 0 : ar.feat[id];
    o.thr[g] = leaf ? 0.f : ar.thr[id];
    o.left[g] = leaf ? 0u : p + 1;
    o.right[g] = rl;
    o.value[g] = leaf ? ar.value[id] : 0u;
    o.nsamp[g] = ar.nsamp[id];
    o.depth[g] = ar.depth[id];
    PNode pn;
    pn.thr = leaf ? 0.f : ar.thr[id];
    pn.feat = leaf ? -1 : ar.feat[id];
    pn.right = leaf ? 0u : base + rl;
    pn.value = leaf ? ar.value[id] : 0u;
    o.pnode[g] = pn;
  }
}

// dst = src with the right-child index of every split node moved by shift
// (a part of a predict table placed after `shift` nodes of another).
__global__ void k_pnode_shift(const PNode *src, uint32_t n, uint32_t shift, PNode *dst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    PNode p = src[i];
    if (p.feat >= 0) p.right += shift;
    dst[i] = p;
  }
}

// End of a level (see ctl_next_body), alone.
__global__ __launch_bounds__(kByteKinds * kWave) void k_ctl_next(Ctl *ctl, unsigned long long *bslot) {
  ctl_next_body(ctl, bslot, threadIdx.x);
}

// Sets the frontier size for a host-provided frontier.
__global__ void k_ctl_set(Ctl *ctl, uint32_t q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->q = q;
}

}  // namespace cta

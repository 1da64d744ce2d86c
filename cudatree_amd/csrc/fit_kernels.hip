// fit_kernels.hip -- level-synchronous random-forest growth on gfx950.
//
// One level of every resident tree is processed by a fixed sequence of
// launches.  Every kernel is wave-granular: a 64-lane wave takes one frontier
// node or one task (a chunk of a node's segment in one or more sorted rows)
// at a time from a grid-stride loop whose bound is read from device memory,
// so the host never has to size grids from device-side counts.
//
//   k_select      per node: feature visiting order, constant-feature skip,
//                 split-scan task list           (bfs_module.cu:122-139)
//   k_scan_hist   per multi-chunk task: class histogram of the chunk
//                                                (dfs_module.cu:24-80)
//   k_scan_carry  per multi-chunk feature: exclusive prefix over chunks
//                                                (dfs_module.cu:82-106)
//   k_scan        per task: Gini scan of the chunk, best position
//                                                (dfs_module.cu:310-443, bfs_module.cu:65-203)
//   k_decide      per node: best split over tasks, threshold, child ids
//                                                (dfs_module.cu:446-491, 158-208, bfs_module.cu:206-235, 379-401)
//   k_fill        per chunk of the split row: mark table, left class counts
//                                                (dfs_module.cu:210-221, bfs_module.cu:238-264)
//   k_children    per split node: leaf tests, next frontier, reshuffle tasks
//                                                (random_tree.py:71-165)
//   k_plan / k_emit / k_resh
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

template <class IdxT, class LabT>
__device__ __forceinline__ const IdxT *rows_of(const FitArgs<IdxT, LabT> &a, uint32_t src, uint32_t tree,
                                               uint32_t f) {
  if (src == 0) return a.si0 + (size_t)f * a.N;
  const IdxT *b = (src == 1) ? a.buf0 : a.buf1;
  return b + ((size_t)tree * a.F + f) * a.N;
}

template <class IdxT, class LabT>
__device__ __forceinline__ IdxT *rows_out(const FitArgs<IdxT, LabT> &a, uint32_t dst, uint32_t tree,
                                          uint32_t f) {
  IdxT *b = (dst == 1) ? a.buf0 : a.buf1;
  return b + ((size_t)tree * a.F + f) * a.N;
}

__device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint32_t smem_wave_offset(int K) { return (threadIdx.x / kWave) * 2u * (uint32_t)K; }

// ------------------------------------------------------------------------
// k_select: visiting order and split-scan tasks of every frontier node.
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_select(FitArgs<IdxT, LabT> a) {
  const uint32_t lane = lane_id();
  const uint32_t m = (uint32_t)a.m, F = (uint32_t)a.F;
  WAVE_LOOP(a.ctl->q) {
    const uint32_t i = item;
    const FNode nd = a.front[i];
    Perm pm = make_perm(a.seed, a.tree_base + nd.tree, nd.depth, nd.start, F, a.identity != 0);
    uint32_t nf = 0;
    for (uint32_t k0 = 0; k0 < F && nf < m; k0 += kWave) {
      const uint32_t k = k0 + lane;
      bool nonconst = false;
      uint32_t f = 0;
      if (k < F) {
        f = perm_at(pm, k);
        const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
        const float *x = a.X_T + (size_t)f * a.N;
        nonconst = x[row[0]] != x[row[nd.n - 1]];
      }
      const uint64_t mk = ballot(nonconst);
      const uint32_t r = mbcnt(mk);
      if (nonconst && nf + r < m) a.feats[(size_t)i * m + nf + r] = (int32_t)f;
      nf += min((uint32_t)__popcll(mk), m - nf);
    }
    const uint32_t n = nd.n;
    const uint32_t nch = ceil_div(n, a.Cs);
    const uint32_t fpt = (nch > 1) ? 1u : max(1u, a.Cs / n);
    const uint32_t ntask = (nf == 0) ? 0u : (nch > 1 ? nf * nch : ceil_div(nf, fpt));
    uint32_t base = 0, hbase = 0;
    if (lane == 0) {
      base = atomicAdd(&a.ctl->n_scan, ntask);
      if (nch > 1) hbase = atomicAdd(&a.ctl->n_hist, nf * nch * (uint32_t)a.K);
      atomicAdd(&a.ctl->bytes_scan, (unsigned long long)nf * n * (sizeof(IdxT) + sizeof(LabT) + 4));
    }
    base = uniform(__shfl(base, 0));
    hbase = uniform(__shfl(hbase, 0));
    const bool fits = (uint64_t)base + ntask <= a.cap_scan &&
                      (nch <= 1 || (uint64_t)hbase + (uint64_t)nf * nch * a.K <= a.cap_hist);
    if (!fits) {
      if (lane == 0) atomicOr(&a.ctl->err, kErrCapacity);
      continue;
    }
    for (uint32_t t = lane; t < ntask; t += kWave) {
      Task tk;
      tk.node = i;
      if (nch > 1) {
        tk.first_count = (t / nch) | (1u << 16);
        tk.chunk = t % nch;
      } else {
        const uint32_t first = t * fpt;
        tk.first_count = first | (min(fpt, nf - first) << 16);
        tk.chunk = 0;
      }
      tk.nch = nch;
      a.scan_tasks[base + t] = tk;
    }
    if (lane == 0) {
      NAux ax = {};
      ax.scan_base = base;
      ax.scan_ntask = ntask;
      ax.nch_s = nch;
      ax.fpt = fpt;
      ax.hist_base = hbase;
      ax.best_f = -1;
      a.aux[i] = ax;
    }
  }
}

// ------------------------------------------------------------------------
// k_scan_hist: class histogram of one chunk of one feature row (multi-chunk
// nodes only).  The prefix over chunks gives each chunk its starting counts,
// the decomposition of scan_gini_large + scan_reduce (dfs_module.cu:24-106).
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_scan_hist(FitArgs<IdxT, LabT> a) {
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
    for (uint32_t k = lane; k < K; k += kWave) h[k] = 0;
    wave_sync();
    const uint32_t p0 = tk.chunk * a.Cs, p1 = min(nd.n, p0 + a.Cs);
    for (uint32_t pb = p0; pb < p1; pb += 4 * kWave) {
      uint32_t sid[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t p = pb + (uint32_t)u * kWave + lane;
        sid[u] = (p < p1) ? (uint32_t)row[p] : 0xFFFFFFFFu;
      }
      uint32_t l[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) l[u] = (sid[u] != 0xFFFFFFFFu) ? (uint32_t)a.y[sid[u]] : 0xFFFFFFFFu;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (l[u] != 0xFFFFFFFFu) atomicAdd(&h[l[u]], 1u);
    }
    wave_sync();
    uint32_t *dst = a.hist + ax.hist_base + ((size_t)j * tk.nch + tk.chunk) * K;
    for (uint32_t k = lane; k < K; k += kWave) dst[k] = h[k];
    wave_sync();
  }
}

// Exclusive prefix over the chunks of each multi-chunk (node, feature).
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_scan_carry(FitArgs<IdxT, LabT> a) {
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  WAVE_LOOP(a.ctl->n_scan) {
    const Task tk = a.scan_tasks[item];
    if (tk.nch <= 1 || tk.chunk != 0) continue;
    const NAux ax = a.aux[tk.node];
    const uint32_t j = tk.first_count & 0xFFFFu;
    uint32_t *base = a.hist + ax.hist_base + (size_t)j * tk.nch * K;
    for (uint32_t k = lane; k < K; k += kWave) {
      uint32_t run = 0;
      for (uint32_t c = 0; c < tk.nch; ++c) {
        uint32_t v = base[(size_t)c * K + k];
        base[(size_t)c * K + k] = run;
        run += v;
      }
    }
  }
}

// ------------------------------------------------------------------------
// k_scan: Gini scan of the task's chunk in each of its features.  A wave
// keeps kScanU 64-position tiles in flight: their index loads, then their
// label and value gathers, are issued together before the tiles are scanned
// in order.
// ------------------------------------------------------------------------
constexpr int kScanU = 4;
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_scan(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  uint32_t *carry = smem + smem_wave_offset(a.K);
  uint32_t *tot = carry + K;
  WAVE_LOOP(a.ctl->n_scan) {
    const Task tk = a.scan_tasks[item];
    const uint32_t i = tk.node;
    const FNode nd = a.front[i];
    const NAux ax = a.aux[i];
    const uint32_t n = nd.n;
    uint64_t t2 = 0;
    for (uint32_t k = lane; k < K; k += kWave) {
      const uint32_t v = a.cnt[(size_t)i * K + k];
      tot[k] = v;
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
      // chunk carry and the running sums at the chunk start
      wave_sync();
      uint64_t sl = 0, dd = 0;
      const uint32_t *hc = (tk.nch > 1) ? a.hist + ax.hist_base + ((size_t)j * tk.nch + tk.chunk) * K : nullptr;
      for (uint32_t k = lane; k < K; k += kWave) {
        const uint32_t c = hc ? hc[k] : 0u;
        carry[k] = c;
        sl += (uint64_t)c * c;
        dd += (uint64_t)tot[k] * c;
      }
      uint64_t SL = uniform64(wave_sum64(sl));
      uint64_t D = uniform64(wave_sum64(dd));
      wave_sync();
      Cand lb = {0, 1, 0, (int32_t)f, 0};
      // value just past the chunk end (the next chunk's first sample)
      const float chunk_next = (p1 < n) ? x[row[p1]] : 0.f;
      for (uint32_t pb = p0; pb < p1; pb += kWave * kScanU) {
        // kScanU tiles in flight: index loads, then label and value gathers
        uint32_t sid[kScanU], lab[kScanU];
        float val[kScanU];
#pragma unroll
        for (int u = 0; u < kScanU; ++u) {
          const uint32_t p = pb + (uint32_t)u * kWave + lane;
          sid[u] = (p < p1) ? (uint32_t)row[p] : 0xFFFFFFFFu;
        }
        const uint32_t pe = pb + kWave * kScanU;  // first position after the batch
        const float batch_next = (pe < p1 && lane == kWave - 1) ? x[row[pe]] : 0.f;
#pragma unroll
        for (int u = 0; u < kScanU; ++u) {
          lab[u] = 0;
          val[u] = 0.f;
          if (sid[u] != 0xFFFFFFFFu) {
            lab[u] = a.y[sid[u]];
            val[u] = x[sid[u]];
          }
        }
#pragma unroll
        for (int u = 0; u < kScanU; ++u) {
          const uint32_t p = pb + (uint32_t)u * kWave + lane;
          const bool valid = p < p1;
          if (!ballot(valid)) break;
          // rank of this lane among earlier lanes with the same label
          uint32_t cprev = 0;
          uint64_t rem = ballot(valid);
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
          const uint32_t dS = valid ? 2u * cprev + 1u : 0u;
          const uint32_t dD = valid ? tot[lab[u]] : 0u;
          const uint32_t iS = wave_incl_scan(dS);
          const uint32_t iD = wave_incl_scan(dD);
          const uint64_t SLp = SL + iS, Dp = D + iD;
          SL += readlane(iS, kWave - 1);
          D += readlane(iD, kWave - 1);
          // value at p + 1: the next lane, the next tile's lane 0, the next
          // batch, or the next chunk
          float nxt = __shfl_down(val[u], 1, kWave);
          if (u + 1 < kScanU) {
            const float first_next = __shfl(val[u + 1 < kScanU ? u + 1 : u], 0, kWave);
            if (lane == kWave - 1) nxt = first_next;
          } else if (lane == kWave - 1) {
            nxt = batch_next;
          }
          if (p + 1 == p1) nxt = chunk_next;
          const bool elig = valid && (p + 1 < n) && gap_ok(val[u], nxt);
          if (elig) {
            const uint64_t nL = (uint64_t)p + 1, nR = (uint64_t)n - nL;
            const uint64_t SR = T2 + SLp - 2 * Dp;
            Cand c = {SLp * nR + SR * nL, nL * nR, p, (int32_t)f, 1u};
            if (!lb.ok || q_cmp(c.num, c.den, lb.num, lb.den) > 0) lb = c;
          }
        }
      }
      lb = wave_best(lb);
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
      a.res[item] = r;
    }
  }
}

// ------------------------------------------------------------------------
// k_decide: best split of each node over its tasks; threshold; child ids.
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_decide(FitArgs<IdxT, LabT> a) {
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
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
    const uint32_t id = nd.nid;
    if (!best.ok) {
      // no candidate position: leaf labelled by the first sample of the
      // segment in feature 0's row (random_tree.py:628-634, 50-63)
      if (lane == 0) {
        const IdxT *row0 = rows_of(a, nd.src, nd.tree, 0) + nd.start;
        a.arena.feat[id] = -1;
        a.arena.thr[id] = 0.f;
        a.arena.left[id] = kLeaf;
        a.arena.right[id] = kLeaf;
        a.arena.value[id] = a.y[row0[0]];
        ax.best_f = -1;
        ax.flags = 0;
        a.aux[i] = ax;
      }
      continue;
    }
    const uint32_t f = (uint32_t)best.feat, p = best.pos;
    for (uint32_t k = lane; k < K; k += kWave) a.cntL[(size_t)i * K + k] = 0;
    if (lane == 0) {
      const IdxT *row = rows_of(a, nd.src, nd.tree, f) + nd.start;
      const float *x = a.X_T + (size_t)f * a.N;
      // threshold midway in float, no contraction (bfs_module.cu:399-400)
      const float thr = __fmul_rn(__fadd_rn(x[row[p]], x[row[p + 1]]), 0.5f);
      const uint32_t ids = atomicAdd(&a.ctl->n_nodes, 2u);
      const uint32_t nch = ceil_div(nd.n, a.Cf);
      const uint32_t fb = atomicAdd(&a.ctl->n_fill, nch);
      if ((uint64_t)ids + 2 > a.cap_nodes || (uint64_t)fb + nch > a.cap_fill) {
        atomicOr(&a.ctl->err, kErrCapacity);
        ax.flags = 0;
        ax.best_f = -1;
      } else {
        a.arena.feat[id] = (int32_t)f;
        a.arena.thr[id] = thr;
        a.arena.left[id] = ids;
        a.arena.right[id] = ids + 1;
        a.arena.value[id] = 0;
        ax.best_f = (int32_t)f;
        ax.best_p = p;
        ax.nL = p + 1;
        ax.left_id = ids;
        ax.right_id = ids + 1;
        ax.flags = kSplit;
        ax.dst = (nd.src == 2) ? 1u : 2u;
        ax.fill_base = fb;
        ax.nch_f = nch;
        for (uint32_t c = 0; c < nch; ++c) a.fill_tasks[fb + c] = Task{i, 0u, c, nch};
        atomicAdd(&a.ctl->bytes_fill, (unsigned long long)nd.n * (sizeof(IdxT) + sizeof(LabT) + 1));
      }
      a.aux[i] = ax;
    }
  }
}

// ------------------------------------------------------------------------
// k_fill: set the bit of every left-going sample of the split row's segment
// in the tree's bitmap (the bitmaps are cleared at the start of the level, so
// right-going samples read 0) and count the left child's classes.
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_fill(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  uint32_t *h = smem + smem_wave_offset(a.K);
  WAVE_LOOP(a.ctl->n_fill) {
    const Task tk = a.fill_tasks[item];
    const FNode nd = a.front[tk.node];
    const NAux ax = a.aux[tk.node];
    const IdxT *row = rows_of(a, nd.src, nd.tree, (uint32_t)ax.best_f) + nd.start;
    uint32_t *bits = a.bits + (size_t)nd.tree * a.W;
    for (uint32_t k = lane; k < K; k += kWave) h[k] = 0;
    wave_sync();
    const uint32_t p0 = tk.chunk * a.Cf, p1 = min(nd.n, p0 + a.Cf);
    for (uint32_t pb = p0; pb < p1; pb += 4 * kWave) {
      uint32_t sid[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t p = pb + (uint32_t)u * kWave + lane;
        sid[u] = (p < p1) ? (uint32_t)row[p] : 0xFFFFFFFFu;
      }
      uint32_t l[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t p = pb + (uint32_t)u * kWave + lane;
        l[u] = (sid[u] != 0xFFFFFFFFu && p <= ax.best_p) ? (uint32_t)a.y[sid[u]] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (l[u] == 0xFFFFFFFFu) continue;  // right side: its bit stays clear
        const uint32_t s = sid[u];
        atomicOr(&bits[s >> 5], 1u << (s & 31u));
        atomicAdd(&h[l[u]], 1u);
      }
    }
    wave_sync();
    for (uint32_t k = lane; k < K; k += kWave)
      if (h[k]) atomicAdd(&a.cntL[(size_t)tk.node * K + k], h[k]);
    wave_sync();
  }
}

__device__ __forceinline__ uint32_t bit_of(const uint32_t *bits, uint32_t s) { return (bits[s >> 5] >> (s & 31u)) & 1u; }

// Reshuffle bookkeeping of split node i (n samples) of tree t: rows of a
// node with at least kBigN samples are moved by whole workgroups (one task
// per row), smaller nodes by single waves (one task per group of rows).
// Counts the node's tasks into its tree's lists; k_plan turns the counts
// into offsets and k_emit writes the task descriptors.
constexpr uint32_t kBigN = 4096;

template <class IdxT, class LabT>
__device__ void count_resh_tasks(const FitArgs<IdxT, LabT> &a, uint32_t t, uint32_t n, NAux &ax) {
  const uint32_t lane = lane_id();
  const uint32_t F = (uint32_t)a.F;
  const bool big = n >= kBigN;
  const uint32_t rpt = big ? 1u : max(1u, a.Cr / n);
  const uint32_t ntask = big ? F : ceil_div(F, rpt);
  uint32_t off = 0;
  if (lane == 0) {
    off = atomicAdd(big ? &a.cnt_big[t] : &a.cnt_small[t], ntask);
    atomicAdd(&a.elems[t], (unsigned long long)F * n);
  }
  off = uniform(__shfl(off, 0));
  ax.resh_big = big ? 1u : 0u;
  ax.resh_off = off;
  ax.resh_ntask = ntask;
  ax.rpt = rpt;
}

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

// ------------------------------------------------------------------------
// k_children: leaf tests of both children (random_tree.py:123-163),
// the next frontier and the reshuffle tasks.
// ------------------------------------------------------------------------
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_children(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  uint32_t *cl = smem + smem_wave_offset(a.K);
  uint32_t *cr = cl + K;
  WAVE_LOOP(a.ctl->q) {
    const uint32_t i = item;
    NAux ax = a.aux[i];
    if (!(ax.flags & kSplit)) continue;
    const FNode nd = a.front[i];
    const uint32_t nL = ax.nL, nR = nd.n - ax.nL;
    uint32_t nzl = 0, nzr = 0, onlyl = 0, onlyr = 0;
    wave_sync();
    for (uint32_t k = lane; k < K; k += kWave) {
      const uint32_t l = a.cntL[(size_t)i * K + k];
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
    uint32_t ol = onlyl, orr = onlyr;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ol = max(ol, (uint32_t)__shfl_xor(ol, o, kWave));
      orr = max(orr, (uint32_t)__shfl_xor(orr, o, kWave));
    }
    wave_sync();
    const uint32_t cd = nd.depth + 1;
    const bool capped = a.max_depth > 0 && (int64_t)cd >= (int64_t)a.max_depth;
    // 0 live, 1 pure, 2 first-sample, 3 majority
    auto kind = [&](uint32_t nz, uint32_t nc) -> uint32_t {
      if (nz <= 1) return 1u;
      if ((int64_t)nc < (int64_t)a.mss) return 2u;
      if (capped) return 3u;
      return 0u;
    };
    const uint32_t kl = kind(nzl, nL), kr = kind(nzr, nR);
    uint32_t labl = ol, labr = orr;
    if (kl == 3) labl = wave_majority(cl, K);
    if (kr == 3) labr = wave_majority(cr, K);
    if (kl == 2 || kr == 2) {
      // first sample of each child in feature 0's row: the first marked
      // (left) and first unmarked (right) sample of the parent's segment
      const IdxT *row0 = rows_of(a, nd.src, nd.tree, 0) + nd.start;
      const uint32_t *bits = a.bits + (size_t)nd.tree * a.W;
      bool needl = kl == 2, needr = kr == 2;
      for (uint32_t pb = 0; pb < nd.n && (needl || needr); pb += kWave) {
        const uint32_t p = pb + lane;
        uint32_t s = 0, side = 2;
        if (p < nd.n) {
          s = row0[p];
          side = bit_of(bits, s);
        }
        const uint64_t ml = ballot(side == 1), mr = ballot(side == 0);
        if (needl && ml) {
          labl = a.y[readlane(s, (uint32_t)__ffsll((unsigned long long)ml) - 1u)];
          needl = false;
        }
        if (needr && mr) {
          labr = a.y[readlane(s, (uint32_t)__ffsll((unsigned long long)mr) - 1u)];
          needr = false;
        }
      }
    }
    uint32_t flags = ax.flags;
    const uint32_t cids[2] = {ax.left_id, ax.right_id};
    const uint32_t kinds[2] = {kl, kr};
    const uint32_t labs[2] = {labl, labr};
    const uint32_t cn[2] = {nL, nR};
    const uint32_t cs[2] = {nd.start, nd.start + nL};
    for (int c = 0; c < 2; ++c) {
      const uint32_t id = cids[c];
      if (lane == 0) {
        a.arena.nsamp[id] = cn[c];
        a.arena.depth[id] = cd;
      }
      if (kinds[c] != 0) {
        if (lane == 0) {
          a.arena.feat[id] = -1;
          a.arena.thr[id] = 0.f;
          a.arena.left[id] = kLeaf;
          a.arena.right[id] = kLeaf;
          a.arena.value[id] = labs[c];
        }
        continue;
      }
      uint32_t slot = 0;
      if (lane == 0) slot = atomicAdd(&a.ctl->q_next, 1u);
      slot = uniform(__shfl(slot, 0));
      if (slot >= a.cap_q) {
        if (lane == 0) atomicOr(&a.ctl->err, kErrCapacity);
        continue;
      }
      if (lane == 0) {
        FNode ch;
        ch.tree = nd.tree;
        ch.start = cs[c];
        ch.n = cn[c];
        ch.nid = id;
        ch.depth = cd;
        ch.src = ax.dst;
        ch.pad0 = ch.pad1 = 0;
        a.next_front[slot] = ch;
      }
      const uint32_t *cc = c ? cr : cl;
      for (uint32_t k = lane; k < K; k += kWave) a.next_cnt[(size_t)slot * K + k] = cc[k];
      flags |= c ? kLiveR : kLiveL;
    }
    ax.flags = flags;
    if (flags & (kLiveL | kLiveR)) {
      count_resh_tasks(a, nd.tree, nd.n, ax);
      if (lane == 0)
        atomicAdd(&a.ctl->bytes_resh, (unsigned long long)a.F * nd.n * (2 * sizeof(IdxT) + 1));
    } else {
      ax.resh_ntask = 0;
    }
    if (lane == 0) a.aux[i] = ax;
    wave_sync();
  }
}

// ------------------------------------------------------------------------
// Reshuffle: stable partition of every row of the split nodes.
//
// k_plan turns the per-tree task counts into offsets and gives each tree
// ceil(elements / kWgElems) workgroups.  k_emit writes the task lists.
// k_resh runs 1,024-thread workgroups, each bound to one tree: it stages the
// tree's side bitmap in LDS (when it fits), then moves its share of the
// tree's big tasks cooperatively (4,096-element tiles, one barrier per tile,
// the left count carried across tiles) and its share of the small tasks one
// wave per task.  Loads and stores are striped (lane-consecutive), and the
// stable ranks come from one ballot per 64 elements.
// ------------------------------------------------------------------------
constexpr int kReshThreads = 1024;
constexpr int kReshWaves = kReshThreads / kWave;  // 16
constexpr int kReshE = 8;                         // elements per thread per tile
constexpr uint32_t kReshTile = kReshThreads * kReshE;
constexpr uint32_t kSlots = kReshE * kReshWaves;  // 128 (sub-tile, wave) counts per tile
static_assert(kSlots == 2 * kWave, "the tile prefix takes two slots per lane");
constexpr uint64_t kWgElems = 1ull << 19;

template <class IdxT, class LabT>
__global__ void k_plan(FitArgs<IdxT, LabT> a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t b = 0, sm = 0, w = 0;
  for (uint32_t t = 0; t < a.T; ++t) {
    a.big_off[t] = b;
    a.small_off[t] = sm;
    a.wg_off[t] = w;
    const unsigned long long e = a.elems[t];
    b += a.cnt_big[t];
    sm += a.cnt_small[t];
    if (e) w += (uint32_t)min((unsigned long long)4096, (e + kWgElems - 1) / kWgElems);
  }
  a.big_off[a.T] = b;
  a.small_off[a.T] = sm;
  a.wg_off[a.T] = w;
  if (b > a.cap_big || sm > a.cap_small) {
    atomicOr(&a.ctl->err, kErrCapacity);
    w = 0;
    a.wg_off[a.T] = 0;
  }
  a.ctl->n_plan = w;
  for (uint32_t t = 0; t < a.T; ++t) {
    a.cnt_big[t] = 0;
    a.cnt_small[t] = 0;
    a.elems[t] = 0;
  }
}

// Writes the task descriptors of every split node with a live child.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_emit(FitArgs<IdxT, LabT> a) {
  const uint32_t lane = lane_id();
  if (a.ctl->n_plan == 0) return;
  WAVE_LOOP(a.ctl->q) {
    const NAux ax = a.aux[item];
    if (!(ax.flags & kSplit) || !(ax.flags & (kLiveL | kLiveR)) || ax.resh_ntask == 0) continue;
    const FNode nd = a.front[item];
    const uint32_t F = (uint32_t)a.F;
    if (ax.resh_big) {
      Task *out = a.big_tasks + a.big_off[nd.tree] + ax.resh_off;
      for (uint32_t f = lane; f < F; f += kWave) out[f] = Task{item, f | (1u << 16), 0u, 1u};
    } else {
      Task *out = a.small_tasks + a.small_off[nd.tree] + ax.resh_off;
      for (uint32_t t = lane; t < ax.resh_ntask; t += kWave) {
        const uint32_t first = t * ax.rpt;
        out[t] = Task{item, first | (min(ax.rpt, F - first) << 16), 0u, 1u};
      }
    }
  }
}

// One big task: all rows' segment [0, n) of one row, moved by the workgroup.
// `par` selects the half of wcnt a tile writes; it alternates across tiles
// and across consecutive tasks, so a tile never overwrites counts that a
// slower wave of the previous tile may still be reading.
template <class IdxT, class LabT, class Side>
__device__ void resh_big(const FitArgs<IdxT, LabT> &a, const Task &tk, const Side &side, uint32_t *wcnt,
                         uint32_t &par) {
  const FNode nd = a.front[tk.node];
  const NAux ax = a.aux[tk.node];
  const uint32_t f = tk.first_count & 0xFFFFu;
  const IdxT *in = rows_of(a, nd.src, nd.tree, f) + nd.start;
  IdxT *out = rows_out(a, ax.dst, nd.tree, f) + nd.start;
  const bool liveL = ax.flags & kLiveL, liveR = ax.flags & kLiveR;
  const uint32_t n = nd.n, nL = ax.nL;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kReshTile) {
    IdxT e[kReshE];
    uint32_t sd[kReshE], rk[kReshE];
#pragma unroll
    for (int k = 0; k < kReshE; ++k) {
      const uint32_t p = base + (uint32_t)k * kReshThreads + tid;
      e[k] = 0;
      if (p < n) e[k] = in[p];
    }
#pragma unroll
    for (int k = 0; k < kReshE; ++k) {
      const uint32_t p = base + (uint32_t)k * kReshThreads + tid;
      sd[k] = (p < n) ? side(e[k]) : 0u;
      const uint64_t m = ballot(sd[k] != 0);
      rk[k] = mbcnt(m);
      if (lane == 0) wcnt[par * kSlots + k * kReshWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    // exclusive prefix over the 128 (sub-tile, wave) counts in position
    // order, two slots per lane
    const uint32_t v0 = wcnt[par * kSlots + 2 * lane], v1 = wcnt[par * kSlots + 2 * lane + 1];
    const uint32_t inc = wave_incl_scan(v0 + v1);
    const uint32_t e0 = inc - v0 - v1, e1 = e0 + v0;
    const uint32_t total = readlane(inc, 63);
    const uint32_t esel = (wave & 1u) ? e1 : e0;  // slot parity = wave parity
#pragma unroll
    for (int k = 0; k < kReshE; ++k) {
      const uint32_t slot = (uint32_t)k * kReshWaves + wave;
      const uint32_t before = carry + readlane(esel, slot >> 1);
      const uint32_t p = base + (uint32_t)k * kReshThreads + tid;
      if (p < n) {
        const uint32_t lb = before + rk[k];
        if (sd[k]) {
          if (liveL) out[lb] = e[k];
        } else if (liveR) {
          out[nL + (p - lb)] = e[k];
        }
      }
    }
    carry += total;
    par ^= 1u;
  }
}

// One small task: rows [first, first+count) of a node, moved by one wave.
template <class IdxT, class LabT, class Side>
__device__ void resh_small(const FitArgs<IdxT, LabT> &a, const Task &tk, const Side &side) {
  const FNode nd = a.front[tk.node];
  const NAux ax = a.aux[tk.node];
  const bool liveL = ax.flags & kLiveL, liveR = ax.flags & kLiveR;
  const uint32_t first = tk.first_count & 0xFFFFu, count = tk.first_count >> 16;
  const uint32_t n = nd.n, nL = ax.nL;
  const uint32_t lane = lane_id();
  for (uint32_t f = first; f < first + count; ++f) {
    const IdxT *in = rows_of(a, nd.src, nd.tree, f) + nd.start;
    IdxT *out = rows_out(a, ax.dst, nd.tree, f) + nd.start;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kWave * kReshE) {
      IdxT e[kReshE];
      uint32_t sd[kReshE];
#pragma unroll
      for (int k = 0; k < kReshE; ++k) {
        const uint32_t p = base + (uint32_t)k * kWave + lane;
        e[k] = 0;
        if (p < n) e[k] = in[p];
      }
#pragma unroll
      for (int k = 0; k < kReshE; ++k) {
        const uint32_t p = base + (uint32_t)k * kWave + lane;
        sd[k] = (p < n) ? side(e[k]) : 0u;
        const uint64_t m = ballot(sd[k] != 0);
        if (p < n) {
          const uint32_t lb = carry + mbcnt(m);
          if (sd[k]) {
            if (liveL) out[lb] = e[k];
          } else if (liveR) {
            out[nL + (p - lb)] = e[k];
          }
        }
        carry += (uint32_t)__popcll(m);
      }
    }
  }
}

// kLds: the tree's bitmap is staged in LDS (N/8 bytes); otherwise it is read
// from global memory.
template <class IdxT, class LabT, bool kLds>
__global__ __launch_bounds__(kReshThreads, 8) void k_resh(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t lds[];
  uint32_t *wcnt = lds + (kLds ? ((a.W + 3u) & ~3u) : 0u);  // 2 x kSlots counts
  const uint32_t P = a.ctl->n_plan;
  const uint32_t T = a.T;
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t loaded = 0xFFFFFFFFu, par = 0;
  for (uint32_t pe = blockIdx.x; pe < P; pe += gridDim.x) {
    // tree of this plan entry: last t with wg_off[t] <= pe
    uint32_t lo = 0, hi = T;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (a.wg_off[mid] <= pe) lo = mid;
      else hi = mid;
    }
    const uint32_t t = lo;
    const uint32_t j = pe - a.wg_off[t], wgs = a.wg_off[t + 1] - a.wg_off[t];
    const uint32_t *gbits = a.bits + (size_t)t * a.W;
    if (kLds && t != loaded) {
      __syncthreads();
      for (uint32_t w = threadIdx.x; w < a.W; w += kReshThreads) lds[w] = gbits[w];
      __syncthreads();
      loaded = t;
    }
    auto side = [&](IdxT s) -> uint32_t {
      const uint32_t u = (uint32_t)s;
      if constexpr (kLds) return (lds[u >> 5] >> (u & 31u)) & 1u;
      else return (gbits[u >> 5] >> (u & 31u)) & 1u;
    };
    const uint32_t nb = a.big_off[t + 1] - a.big_off[t];
    for (uint32_t bt = j; bt < nb; bt += wgs) resh_big(a, a.big_tasks[a.big_off[t] + bt], side, wcnt, par);
    const uint32_t ns = a.small_off[t + 1] - a.small_off[t];
    for (uint32_t st = j * kReshWaves + wave; st < ns; st += wgs * kReshWaves)
      resh_small(a, a.small_tasks[a.small_off[t] + st], side);
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
// (src 0, N samples) keep the in-bag samples (mark 1) in buffer 0.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_boot_tasks(FitArgs<IdxT, LabT> a, const uint64_t *off, uint32_t T,
                                                       FNode *front, uint32_t dummy) {
  const uint32_t lane = lane_id();
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
      nd.pad0 = nd.pad1 = dummy;
      front[t] = nd;
    }
    NAux ax = {};
    ax.best_f = -1;
    ax.nL = (uint32_t)(off[t + 1] - off[t]);
    ax.flags = kSplit | kLiveL;
    ax.dst = 1;
    count_resh_tasks(a, t, (uint32_t)a.N, ax);
    if (lane == 0) {
      a.aux[t] = ax;
      atomicAdd(&a.ctl->bytes_boot, (unsigned long long)a.F * a.N * (sizeof(IdxT) + 1) +
                                        (unsigned long long)a.F * ax.nL * sizeof(IdxT));
    }
  }
}

// Roots: tree t's root (arena id t) is a leaf if pure, else enters the frontier.
template <class IdxT, class LabT>
__global__ __launch_bounds__(kBlock) void k_roots(FitArgs<IdxT, LabT> a, const uint32_t *root_cnt,
                                                  const uint64_t *off, uint32_t T, uint32_t src) {
  const uint32_t lane = lane_id();
  const uint32_t K = (uint32_t)a.K;
  WAVE_LOOP(T) {
    const uint32_t t = item;
    const uint32_t n = off ? (uint32_t)(off[t + 1] - off[t]) : (uint32_t)a.N;
    uint32_t nz = 0, only = 0;
    for (uint32_t k = lane; k < K; k += kWave)
      if (root_cnt[(size_t)t * K + k]) {
        ++nz;
        only = k;
      }
    nz = uniform(wave_sum32(nz));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) only = max(only, (uint32_t)__shfl_xor(only, o, kWave));
    if (lane == 0) {
      a.arena.nsamp[t] = n;
      a.arena.depth[t] = 0;
    }
    // The reference never tests the root for purity (random_tree.py:482-486);
    // a pure root is a leaf here.  min_samples_split applies as at any node;
    // a positive max_depth never caps the root.
    uint32_t leaf_label = kLeaf;
    if (nz <= 1) leaf_label = only;
    else if ((int64_t)n < (int64_t)a.mss) {
      const IdxT *row0 = rows_of(a, src, t, 0);
      leaf_label = a.y[row0[0]];
    }
    if (leaf_label != kLeaf) {
      if (lane == 0) {
        a.arena.feat[t] = -1;
        a.arena.thr[t] = 0.f;
        a.arena.left[t] = kLeaf;
        a.arena.right[t] = kLeaf;
        a.arena.value[t] = leaf_label;
      }
      continue;
    }
    uint32_t slot = 0;
    if (lane == 0) slot = atomicAdd(&a.ctl->q_next, 1u);
    slot = uniform(__shfl(slot, 0));
    if (lane == 0) {
      FNode nd;
      nd.tree = t;
      nd.start = 0;
      nd.n = n;
      nd.nid = t;
      nd.depth = 0;
      nd.src = src;
      nd.pad0 = nd.pad1 = 0;
      a.next_front[slot] = nd;
    }
    for (uint32_t k = lane; k < K; k += kWave) a.next_cnt[(size_t)slot * K + k] = root_cnt[(size_t)t * K + k];
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

// Moves the next frontier's count into place and clears the level counters.
__global__ void k_ctl_next(Ctl *ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->q = ctl->q_next;
    ctl->q_next = 0;
    ctl->n_scan = ctl->n_fill = ctl->n_hist = ctl->n_plan = 0;
  }
}

// Sets the frontier size for a host-provided frontier.
__global__ void k_ctl_set(Ctl *ctl, uint32_t q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->q = q;
    ctl->q_next = 0;
    ctl->n_scan = ctl->n_fill = ctl->n_hist = ctl->n_plan = 0;
  }
}

}  // namespace cta

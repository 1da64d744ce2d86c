// subtree_kernels.hip -- small nodes grown to their leaves by single waves.
//
// The levels (fit_kernels.hip) process every node of a level in a fixed
// sequence of launches, which is what large nodes need; below sub_max
// samples a node's whole subtree is cheap work, but level by level it would
// pay that sequence at every remaining depth (the deep levels of a forest
// hold nodes of a few hundred samples).  Live children of at most sub_max
// samples therefore leave the levels for a queue, and one launch of
// k_subtree grows every queued node of every tree: each wave takes queue
// entries and finishes one subtree at a time depth first, with the node
// steps of the levels done in place by the wave itself -- class counts,
// feature visiting order with the constant-feature skip, exact Gini scans
// of the rows, the split choice and threshold, the children's leaf tests,
// the side marks and the stable partition of every row into the other
// buffer (random_tree.py:580-676, dfs_module.cu, bfs_module.cu).  A node's
// segment stays at the same offsets in every row, so subtrees of one tree
// never touch each other's rows, and only the marks of the tree's side
// bitmap are shared: they are set and cleared with atomics, each subtree on
// its own samples.  The smaller child is taken first, so a wave's stack of
// pending nodes never holds more than log2(sub_max) + 2 of them.
#include "device_util.h"
#include "internal.h"

namespace cta {

constexpr int kSubThreads = 256;
constexpr int kSubWaves = kSubThreads / kWave;
// Pending nodes of a wave, one per lane (start, n, id, depth, rows).
constexpr uint32_t kSubStack = 40;

// Per-wave LDS words: class totals, class carries (split scans with K > 64
// or LDS counters), left-child counts and the row-free counters.
__host__ __device__ __forceinline__ uint32_t sub_wave_words(uint32_t K, uint32_t Frl) { return 3 * K + Frl * K; }

// The side-bitmap word holding sample s, read from the device-coherent copy
// (the marks are set by atomics of this wave).
__device__ __forceinline__ uint32_t mark_word(const uint32_t *bits, uint32_t s) {
  return __hip_atomic_load(bits + (s >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sets (set = true) or clears the marks of the samples at positions
// [p0, p1) of row `row`, or of those of them whose bit in column `col` (a
// row-free feature's value bits; nullptr: every one) equals `want`.
template <class IdxT>
__device__ __forceinline__ void sub_marks(const IdxT *row, uint32_t p0, uint32_t p1, const uint32_t *col,
                                          uint32_t want, uint32_t *bits, bool set) {
  const uint32_t lane = lane_id();
  for (uint32_t p = p0 + lane; p < p1; p += kWave) {
    const uint32_t s = sample_of(row[p]);
    if (col && ((col[s >> 5] >> (s & 31u)) & 1u) != want) continue;
    const uint32_t b = 1u << (s & 31u);
    if (set) atomicOr(&bits[s >> 5], b);
    else atomicAnd(&bits[s >> 5], ~b);
  }
}

template <class IdxT, class LabT>
__global__ __launch_bounds__(kSubThreads) void k_subtree(FitArgs<IdxT, LabT> a) {
  extern __shared__ uint32_t smem[];
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t K = (uint32_t)a.K, F = (uint32_t)a.F, m = (uint32_t)a.m, Frl = a.Frl, FK = a.Frl * (uint32_t)a.K;
  uint32_t *tot = smem + wave * sub_wave_words(K, Frl);
  uint32_t *carry = tot + K, *cl = carry + K, *rl = cl + K;
  const bool regs = K <= (uint32_t)kWave && !a.hist_lds;
  const uint32_t nq = a.ctl->n_sub;
  unsigned long long bytes = 0;
  uint32_t max_depth_seen = 0;
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(&a.ctl->sub_next, 1u);
    item = readlane(item, 0);
    if (item >= nq || a.ctl->err) break;
    const FNode root = a.sub_queue[item];
    const uint32_t tree = uniform(root.tree);
    uint32_t *bits = a.bits + (size_t)tree * a.W;
    // the pending nodes: lane j holds entry j
    uint32_t s_start = 0, s_n = 0, s_id = 0, s_depth = 0, s_src = 0;
    if (lane == 0) {
      s_start = root.start;
      s_n = root.n;
      s_id = root.nid;
      s_depth = root.depth;
      s_src = root.src;
    }
    uint32_t sp = 1;
    while (sp > 0) {
      --sp;
      const uint32_t start = readlane(s_start, sp), n = readlane(s_n, sp), nid = readlane(s_id, sp);
      const uint32_t depth = readlane(s_depth, sp), src = readlane(s_src, sp);
      const uint32_t dst = src == 1 ? 2u : 1u;
      const IdxT *row0 = rows_of(a, src, tree, 0) + start;

      // ---- class totals and row-free counts over the feature-0 row
      uint32_t tv = 0;  // regs: lane k holds class k
      for (uint32_t k = lane; k < K; k += kWave) tot[k] = 0;
      for (uint32_t w = lane; w < FK; w += kWave) rl[w] = 0;
      wave_sync();
      for (uint32_t pb = 0; pb < n; pb += 4 * kWave) {
        uint32_t e[4], l[4];
        bool take[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t p = pb + (uint32_t)u * kWave + lane;
          take[u] = p < n;
          e[u] = take[u] ? (uint32_t)row0[p] : 0u;
          l[u] = take[u] ? a.label((IdxT)e[u]) : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (regs) {
            uint64_t rem = ballot(take[u]);
            while (rem) {
              const uint32_t L = readlane(l[u], (uint32_t)__ffsll((unsigned long long)rem) - 1u);
              const uint64_t mm = ballot(take[u] && l[u] == L);
              if (lane == L) tv += (uint32_t)__popcll(mm);
              rem &= ~mm;
            }
          } else if (take[u]) {
            atomicAdd(&tot[l[u]], 1u);
          }
        }
        if (Frl) rl_add(a, e, take, rl);
      }
      wave_sync();
      if (regs && lane < K) tot[lane] = tv;  // every lane reads the totals of the row-free candidates
      wave_sync();
      uint64_t t2 = 0;
      for (uint32_t k = lane; k < K; k += kWave) t2 += (uint64_t)tot[k] * tot[k];
      const uint64_t T2 = uniform64(wave_sum64(t2));

      // ---- features in visiting order: constants skipped until m are
      // evaluated; row-free ones scored from their counts, the others scanned
      const Perm pm = make_perm(a.seed, a.tree_base + tree, depth, start, F, a.identity != 0);
      Cand best = {0, 1, 0, 0, 0}, rlb = {0, 1, 0, -1, 0};
      uint32_t nsel = 0, nf = 0;
      for (uint32_t k0 = 0; k0 < F && nsel < m; k0 += kWave) {
        const uint32_t k = k0 + lane;
        bool nonconst = false;
        uint32_t f = 0;
        int32_t jr = -1;
        Cand cd = {0, 1, 0, -1, 0};
        if (k < F) {
          f = perm_at(pm, k);
          jr = Frl ? a.rl_of[f] : -1;
          if (jr >= 0) {
            const uint32_t *oc = rl + (uint32_t)jr * K;
            uint64_t SL = 0, Dd = 0;
            uint32_t ones = 0;
            for (uint32_t c = 0; c < K; ++c) {
              const uint64_t t = tot[c], o = oc[c], cl0 = t - o;
              SL += cl0 * cl0;
              Dd += t * cl0;
              ones += (uint32_t)o;
            }
            nonconst = ones != 0u && ones != n;
            const uint64_t nL = n - ones, nR = ones;
            const uint64_t SR = T2 + SL - 2 * Dd;
            cd = {SL * nR + SR * nL, nL * nR, (uint32_t)nL - 1u, (int32_t)f, 1u};
          } else {
            const IdxT *row = rows_of(a, src, tree, f) + start;
            const uint32_t s0 = sample_of(row[0]), s1 = sample_of(row[n - 1]);
            const uint64_t fc = a.fcode[f];
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
        if (take && jr >= 0 && cand_better(cd, rlb)) rlb = cd;
        uint64_t mrow = ballot(take && jr < 0);
        nsel += min((uint32_t)__popcll(mk), m - nsel);
        while (mrow) {
          const uint32_t fr = readlane(f, (uint32_t)__ffsll((unsigned long long)mrow) - 1u);
          mrow &= mrow - 1u;
          ++nf;
          const IdxT *row = rows_of(a, src, tree, fr) + start;
          const float *x = a.X_T + (size_t)fr * a.N;
          const uint64_t fc = a.fcode[fr];
          const uint32_t kind = uniform(fc_kind(fc));
          const uint8_t *cb = a.codes + fc_off(fc);
          auto vat = [&](uint32_t s) -> uint32_t { return kind ? code_at(cb, kind, s) : __float_as_uint(x[s]); };
          auto cand = [&](uint32_t v, uint32_t w) -> bool {
            return kind ? v != w : gap_ok(__uint_as_float(v), __uint_as_float(w));
          };
          for (uint32_t c = lane; c < K; c += kWave) carry[c] = 0;
          wave_sync();
          const Cand c = regs ? scan_positions<true>(a, row, 0u, n, n, fr, 0u, tv, carry, tot, 0ull, 0ull, T2, 0u,
                                                     vat, cand)
                              : scan_positions<false>(a, row, 0u, n, n, fr, 0u, tv, carry, tot, 0ull, 0ull, T2, 0u,
                                                      vat, cand);
          if (cand_better(c, best)) best = c;
        }
      }
      rlb = wave_best(rlb);
      const bool rowfree = cand_better(rlb, best);
      if (rowfree) best = rlb;
      bytes += (unsigned long long)n * (sizeof(IdxT) + sizeof(LabT) + 4 * a.Wsig) +
               (unsigned long long)nf * n * (sizeof(IdxT) + sizeof(LabT) + 4);

      if (!best.ok) {
        // no candidate position: leaf labelled by the segment's first sample
        if (lane == 0) {
          a.arena.feat[nid] = -1;
          a.arena.thr[nid] = 0.f;
          a.arena.left[nid] = kLeaf;
          a.arena.right[nid] = kLeaf;
          a.arena.value[nid] = a.label(row0[0]);
        }
        continue;
      }

      // ---- the split: threshold and the left child's class counts
      const uint32_t f = (uint32_t)best.feat, p = best.pos, nL = p + 1, nR = n - nL;
      float thr;
      const IdxT *srow = nullptr;  // the split feature's row (row splits)
      const uint32_t *col = nullptr;  // the split feature's value bits (row-free splits)
      if (rowfree) {
        const uint32_t jr = (uint32_t)a.rl_of[f];
        thr = a.rl_thr[jr];
        col = reinterpret_cast<const uint32_t *>(a.codes + fc_off(a.fcode[f]));
        for (uint32_t c = lane; c < K; c += kWave) cl[c] = tot[c] - rl[jr * K + c];
      } else {
        srow = rows_of(a, src, tree, f) + start;
        const float *x = a.X_T + (size_t)f * a.N;
        // midway in float, no contraction (bfs_module.cu:399-400)
        thr = __fmul_rn(__fadd_rn(x[sample_of(srow[p])], x[sample_of(srow[p + 1])]), 0.5f);
        for (uint32_t c = lane; c < K; c += kWave) cl[c] = 0;
        wave_sync();
        for (uint32_t pb = 0; pb < nL; pb += 4 * kWave) {
          uint32_t l[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t q = pb + (uint32_t)u * kWave + lane;
            l[u] = q < nL ? a.label(srow[q]) : 0xFFFFFFFFu;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (l[u] != 0xFFFFFFFFu) atomicAdd(&cl[l[u]], 1u);
        }
      }
      wave_sync();
      // right-child counts into carry; present classes of each child
      uint32_t nzl = 0, nzr = 0, onlyl = 0, onlyr = 0;
      for (uint32_t c = lane; c < K; c += kWave) {
        const uint32_t l = cl[c], r = tot[c] - l;
        carry[c] = r;
        if (l) {
          ++nzl;
          onlyl = c;
        }
        if (r) {
          ++nzr;
          onlyr = c;
        }
      }
      nzl = uniform(wave_sum32(nzl));
      nzr = uniform(wave_sum32(nzr));
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        onlyl = max(onlyl, (uint32_t)__shfl_xor(onlyl, o, kWave));
        onlyr = max(onlyr, (uint32_t)__shfl_xor(onlyr, o, kWave));
      }
      wave_sync();
      const uint32_t kl = child_kind(a, nzl, nL, depth), kr = child_kind(a, nzr, nR, depth);
      uint32_t labl = uniform(onlyl), labr = uniform(onlyr);
      if (kl == 3) labl = wave_majority(cl, K);
      if (kr == 3) labr = wave_majority(carry, K);

      // ---- node ids of the children
      uint32_t id0 = 0;
      if (lane == 0) id0 = atomicAdd(&a.ctl->n_nodes, 2u);
      id0 = readlane(id0, 0);
      if ((uint64_t)id0 + 2 > a.cap_nodes || sp + 2 > kSubStack) {
        if (lane == 0) atomicOr(&a.ctl->err, kErrCapacity);
        break;
      }

      // ---- side marks of the smaller child, the first samples of children
      // labelled by them, and the partition of every row
      const bool small_left = nL <= nR;
      const uint32_t inv = small_left ? 0u : 1u;  // marks on the right child
      const bool part = kl == 0 || kr == 0;
      if (part || kl == 2 || kr == 2) {
        if (rowfree) sub_marks(row0, 0u, n, col, small_left ? 0u : 1u, bits, true);  // lower value left
        else sub_marks(srow, small_left ? 0u : nL, small_left ? nL : n, (const uint32_t *)nullptr, 0u, bits, true);
        __threadfence();
        if (kl == 2 || kr == 2) {
          // first marked (left) and unmarked (right) samples of feature 0's row
          bool needl = kl == 2, needr = kr == 2;
          for (uint32_t pb = 0; pb < n && (needl || needr); pb += kWave) {
            const uint32_t q = pb + lane;
            uint32_t e = 0, side = 2;
            if (q < n) {
              e = (uint32_t)row0[q];
              const uint32_t s = sample_of((IdxT)e);
              side = ((mark_word(bits, s) >> (s & 31u)) & 1u) ^ inv;
            }
            const uint64_t ml = ballot(side == 1), mr = ballot(side == 0);
            if (needl && ml) {
              labl = a.label((IdxT)readlane(e, (uint32_t)__ffsll((unsigned long long)ml) - 1u));
              needl = false;
            }
            if (needr && mr) {
              labr = a.label((IdxT)readlane(e, (uint32_t)__ffsll((unsigned long long)mr) - 1u));
              needr = false;
            }
          }
        }
        if (part) {
          const uint32_t flags = (kl == 0 ? kLiveL : 0u) | (kr == 0 ? kLiveR : 0u) | (inv ? kInvert : 0u);
          auto side = [&](IdxT e) -> uint32_t {
            const uint32_t s = sample_of(e);
            return (mark_word(bits, s) >> (s & 31u)) & 1u;
          };
          resh_rows(a, tree, start, src, dst, n, nL, flags, 0u, a.Fr, side);
          bytes += (unsigned long long)a.Fr * n * (2 * sizeof(IdxT) + 1);
        }
        if (rowfree) sub_marks(row0, 0u, n, col, small_left ? 0u : 1u, bits, false);
        else sub_marks(srow, small_left ? 0u : nL, small_left ? nL : n, (const uint32_t *)nullptr, 0u, bits, false);
        __threadfence();
        bytes += (unsigned long long)(rowfree ? n : min(nL, nR)) * (sizeof(IdxT) + 1);
      }

      // ---- the node and its children in the arena; live children pending,
      // the larger first so the smaller is taken next
      if (lane == 0) {
        a.arena.feat[nid] = (int32_t)f;
        a.arena.thr[nid] = thr;
        a.arena.left[nid] = id0;
        a.arena.right[nid] = id0 + 1;
        a.arena.value[nid] = 0;
        for (uint32_t c = 0; c < 2; ++c) {
          const uint32_t id = id0 + c, kc = c ? kr : kl;
          a.arena.nsamp[id] = c ? nR : nL;
          a.arena.depth[id] = depth + 1;
          if (kc) {
            a.arena.feat[id] = -1;
            a.arena.thr[id] = 0.f;
            a.arena.left[id] = kLeaf;
            a.arena.right[id] = kLeaf;
            a.arena.value[id] = c ? labr : labl;
          }
        }
      }
      max_depth_seen = max(max_depth_seen, depth + 1);
      const uint32_t order[2] = {small_left ? 1u : 0u, small_left ? 0u : 1u};  // larger, then smaller
      for (uint32_t j = 0; j < 2; ++j) {
        const uint32_t c = order[j];
        if ((c ? kr : kl) != 0) continue;
        if (lane == sp) {
          s_start = c ? start + nL : start;
          s_n = c ? nR : nL;
          s_id = id0 + c;
          s_depth = depth + 1;
          s_src = dst;
        }
        ++sp;
      }
    }
  }
  if (lane == 0) {
    if (bytes) atomicAdd(&a.ctl->bytes_sub, bytes);
    if (max_depth_seen) atomicMax(&a.ctl->sub_depth, max_depth_seen);
  }
}

}  // namespace cta

// aux_kernels.hip -- training-set preparation and prediction on gfx950.
//
//   k_transpose        X (N x F, row-major) -> X_T (F x N) through LDS tiles
//                      (random_forest.py:233)
//   k_sort_*           stable LSD radix sort of every feature row, 4-bit
//                      digits; a feature takes only the passes of digits
//                      that vary over its keys (one pass when it has two
//                      distinct keys); ties keep ascending sample id
//                      (replaces the host np.argsort of
//                      random_forest.py:247)
//   k_block_*          blocked predict table (subtrees of three levels
//                      per 128-byte line) from the pre-order table
//   k_predict / k_vote batched traversal of the blocked forest, one thread
//                      per (row, tree), votes by integer atomics
//                      (predict.cu:12-69, random_forest.py:339-377)
#include "device_util.h"
#include "internal.h"

namespace cta {

// ---- transpose ------------------------------------------------------------
constexpr int kTT = 64;

__global__ __launch_bounds__(256) void k_transpose(const float *__restrict__ X, uint32_t N, uint32_t F,
                                                   float *__restrict__ XT, uint32_t *err) {
  __shared__ float tile[kTT][kTT + 1];
  const uint32_t r0 = blockIdx.x * kTT, f0 = blockIdx.y * kTT;
  const uint32_t tx = threadIdx.x % kTT, ty = threadIdx.x / kTT;  // 64 x 4
  bool nan = false;
  for (uint32_t j = ty; j < kTT; j += 4) {
    const uint32_t r = r0 + j, f = f0 + tx;
    float v = 0.f;
    if (r < N && f < F) {
      v = X[(size_t)r * F + f];
      nan |= v != v;
    }
    tile[j][tx] = v;
  }
  __syncthreads();
  for (uint32_t j = ty; j < kTT; j += 4) {
    const uint32_t f = f0 + j, r = r0 + tx;
    if (r < N && f < F) XT[(size_t)f * N + r] = tile[tx][j];
  }
  if (nan) atomicOr(err, 1u);
}

// ---- stable radix presort -------------------------------------------------
constexpr int kSortThreads = 256;
#ifndef CTA_SORT_ITEMS
#define CTA_SORT_ITEMS 8
#endif
constexpr int kSortItems = CTA_SORT_ITEMS;
constexpr int kSortTile = kSortThreads * kSortItems;  // elements per block and pass
constexpr int kDigits = 16;

// Order-preserving key of a float: -0.0 and +0.0 map to the same key so that
// they tie (and keep id order) as in a comparison sort.
__device__ __forceinline__ uint32_t float_key(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7FFFFFFFu) == 0) u = 0;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Per feature f, over its keys: mask[f] |= OR, mask[F + f] &= AND (the bits
// that vary over the feature are their XOR), mask[2F + f] = min and
// mask[3F + f] = max (set to 0, ~0, ~0 and 0 first).
__global__ __launch_bounds__(256) void k_key_mask(const float *XT, uint32_t N, uint32_t F, uint32_t *mask) {
  const uint32_t f = blockIdx.y;
  uint32_t o = 0u, n = ~0u, lo = ~0u, hi = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const uint32_t k = float_key(XT[(size_t)f * N + i]);
    o |= k;
    n &= k;
    lo = min(lo, k);
    hi = max(hi, k);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    o |= (uint32_t)__shfl_xor((int)o, d, 64);
    n &= (uint32_t)__shfl_xor((int)n, d, 64);
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, d, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, d, 64));
  }
  if (lane_id() == 0) {
    atomicOr(&mask[f], o);
    atomicAnd(&mask[F + f], n);
    atomicMin(&mask[2 * F + f], lo);
    atomicMax(&mask[3 * F + f], hi);
  }
}

// mask[4F + f] += keys of feature f other than its min and max (set to 0
// first): zero for a feature with at most two distinct keys.
__global__ __launch_bounds__(256) void k_key_others(const float *XT, uint32_t N, uint32_t F, uint32_t *mask) {
  const uint32_t f = blockIdx.y;
  const uint32_t lo = mask[2 * F + f], hi = mask[3 * F + f];
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const uint32_t k = float_key(XT[(size_t)f * N + i]);
    c += (k != lo && k != hi) ? 1u : 0u;
  }
  c = wave_sum32(c);
  if (lane_id() == 0 && c) atomicAdd(&mask[4 * F + f], c);
}

// Keys and ids of every feature, written into buffer 1 for the features
// with an odd number of passes (odd[f] != 0), else buffer 0, so that every
// feature ends its passes in buffer 0.
__global__ void k_sort_init(const float *XT, uint32_t N, uint32_t F, const uint32_t *odd, uint32_t *k0,
                            uint32_t *v0, uint32_t *k1, uint32_t *v1) {
  const size_t total = (size_t)N * F;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const bool b1 = odd[i / N] != 0u;
    (b1 ? k1 : k0)[i] = float_key(XT[i]);
    (b1 ? v1 : v0)[i] = (uint32_t)(i % N);
  }
}

// The features of one pass: entry j of `list` is feature f, with bit 31
// set when its keys are in buffer 1 before the pass.
struct SortPass {
  const uint32_t *list;
  uint32_t *k0, *v0, *k1, *v1;
};

// hist[(f * 16 + d) * ntiles + tile] = count of digit d in the tile
// (feature f = entry blockIdx.y of the pass's list).
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(SortPass sp, uint32_t N, uint32_t ntiles, uint32_t shift,
                                                            uint32_t *hist) {
  __shared__ uint32_t h[kDigits];
  const uint32_t tile = blockIdx.x, ent = sp.list[blockIdx.y], f = ent & 0x7FFFFFFFu;
  if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t *k = ((ent >> 31) ? sp.k1 : sp.k0) + (size_t)f * N;
  const uint32_t base = tile * kSortTile;
  for (uint32_t j = threadIdx.x; j < (uint32_t)kSortTile; j += kSortThreads) {
    const uint32_t e = base + j;
    if (e < N) atomicAdd(&h[(k[e] >> shift) & (kDigits - 1)], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kDigits) hist[((size_t)f * kDigits + threadIdx.x) * ntiles + tile] = h[threadIdx.x];
}

// Exclusive scan of each listed row's 16 x ntiles counts (digit-major): the
// output offset of digit d's elements of each tile.
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(SortPass sp, uint32_t *hist, uint32_t ntiles) {
  __shared__ uint32_t part[kSortThreads];
  uint32_t *h = hist + (size_t)(sp.list[blockIdx.x] & 0x7FFFFFFFu) * kDigits * ntiles;
  const uint32_t len = kDigits * ntiles;
  const uint32_t per = (len + kSortThreads - 1) / kSortThreads;
  const uint32_t a = threadIdx.x * per, b = min(len, a + per);
  uint32_t s = 0;
  for (uint32_t i = a; i < b; ++i) s += h[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int t = 0; t < kSortThreads; ++t) {
      uint32_t v = part[t];
      part[t] = run;
      run += v;
    }
  }
  __syncthreads();
  uint32_t run = part[threadIdx.x];
  for (uint32_t i = a; i < b; ++i) {
    uint32_t v = h[i];
    h[i] = run;
    run += v;
  }
}

// Stable scatter of one tile: rounds of 256 consecutive elements, ranks by
// one ballot per digit inside each wave, wave and round offsets from LDS.
// (from the feature's buffer into the other one)
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(SortPass sp, uint32_t N, uint32_t ntiles,
                                                               uint32_t shift, const uint32_t *hist) {
  __shared__ uint32_t wcnt[kSortThreads / 64][kDigits];
  __shared__ uint32_t run[kDigits];
  __shared__ uint32_t off[kDigits];
  const uint32_t tile = blockIdx.x, ent = sp.list[blockIdx.y], f = ent & 0x7FFFFFFFu;
  const bool in1 = ent >> 31;
  const uint32_t *kin = in1 ? sp.k1 : sp.k0, *vin = in1 ? sp.v1 : sp.v0;
  uint32_t *kout = in1 ? sp.k0 : sp.k1, *vout = in1 ? sp.v0 : sp.v1;
  const uint32_t lane = lane_id(), w = threadIdx.x / 64;
  if (threadIdx.x < kDigits) {
    run[threadIdx.x] = 0;
    off[threadIdx.x] = hist[((size_t)f * kDigits + threadIdx.x) * ntiles + tile];
  }
  __syncthreads();
  const size_t rb = (size_t)f * N;
  for (uint32_t r = 0; r < (uint32_t)kSortItems; ++r) {
    const uint32_t e = tile * kSortTile + r * kSortThreads + threadIdx.x;
    const bool valid = e < N;
    uint32_t key = 0, val = 0, d = kDigits;
    if (valid) {
      key = kin[rb + e];
      val = vin[rb + e];
      d = (key >> shift) & (kDigits - 1);
    }
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t dd = 0; dd < (uint32_t)kDigits; ++dd) {
      const uint64_t m = ballot(d == dd);
      if (d == dd) rank = mbcnt(m);
      if (lane == 0) wcnt[w][dd] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = off[d] + run[d] + rank;
      for (uint32_t ww = 0; ww < w; ++ww) pos += wcnt[ww][d];
      kout[rb + pos] = key;
      vout[rb + pos] = val;
    }
    __syncthreads();
    if (threadIdx.x < kDigits) {
      uint32_t s = 0;
      for (int ww = 0; ww < kSortThreads / 64; ++ww) s += wcnt[ww][threadIdx.x];
      run[threadIdx.x] += s;
    }
    __syncthreads();
  }
}

// Sorted ids to row entries: 16-bit ids, or 32-bit words with the label in
// the high bits.
template <class IdxT, class LabT>
__global__ void k_sort_finish(const uint32_t *vals, size_t total, const LabT *y, IdxT *out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t s = vals[i];
    if constexpr (sizeof(IdxT) == 4) out[i] = (IdxT)(s | ((uint32_t)y[s] << kIdBits));
    else out[i] = (IdxT)s;
  }
}

// Row entries to plain 32-bit sample ids.
template <class IdxT>
__global__ void k_widen(const IdxT *in, size_t total, uint32_t *out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
    out[i] = sample_of(in[i]);
}

template <class LabT>
__global__ void k_labels(const uint32_t *y, uint32_t N, uint32_t K, LabT *out, uint32_t *err) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const uint32_t v = y[i];
    if (v >= K) atomicOr(err, 2u);
    out[i] = (LabT)(v < K ? v : 0);
  }
}

// ---- label compaction ---------------------------------------------------------
// Class values arrive as order-preserving 64-bit keys; their distinct keys in
// ascending order are the classes, and a label is its key's index among them
// (np.unique(..., return_inverse=True), random_forest.py:85-99).
// k_key_distinct reduces a list of keys to the distinct keys of each tile of
// kKeyTile (bitonic sort in LDS, then the first of every run), appended to an
// output list in any order; repeated until one tile holds the list, whose
// sorted distinct keys k_key_table writes.  k_key_map: each label by binary
// search in the table (held in LDS).
constexpr int kKeyThreads = 256;
constexpr uint32_t kKeyTile = 2048;

// Sorts the kKeyTile keys in LDS ascending (bitonic network; all threads).
__device__ __forceinline__ void lds_bitonic(unsigned long long *k) {
  for (uint32_t size = 2; size <= kKeyTile; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (uint32_t t = threadIdx.x; t < kKeyTile / 2; t += kKeyThreads) {
        const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned long long a = k[lo], b = k[hi];
        if ((a > b) == up) {
          k[lo] = b;
          k[hi] = a;
        }
      }
    }
  __syncthreads();
}

// Tile b of keys[0, n): its keys sorted, padded with copies of its first key.
__device__ __forceinline__ void load_key_tile(const unsigned long long *keys, uint32_t n, uint32_t b,
                                              unsigned long long *k) {
  const uint32_t base = b * kKeyTile;
  const unsigned long long pad = keys[base];
  for (uint32_t j = threadIdx.x; j < kKeyTile; j += kKeyThreads) k[j] = base + j < n ? keys[base + j] : pad;
  lds_bitonic(k);
}

__global__ __launch_bounds__(kKeyThreads) void k_key_distinct(const unsigned long long *keys, uint32_t n,
                                                              unsigned long long *out, uint32_t *count) {
  __shared__ unsigned long long k[kKeyTile];
  __shared__ uint32_t base;
  if (blockIdx.x * kKeyTile >= n) return;
  load_key_tile(keys, n, blockIdx.x, k);
  // thread t owns the contiguous positions [t E, t E + E), so the distinct
  // keys leave in sorted order
  constexpr uint32_t E = kKeyTile / kKeyThreads;
  const uint32_t j0 = threadIdx.x * E;
  uint32_t nd = 0;
  for (uint32_t j = j0; j < j0 + E; ++j) nd += (j == 0 || k[j] != k[j - 1]) ? 1u : 0u;
  // distinct keys of the tile, appended in tile order at one reserved range
  __shared__ uint32_t part[kKeyThreads];
  part[threadIdx.x] = nd;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kKeyThreads; ++i) {
      const uint32_t v = part[i];
      part[i] = t;
      t += v;
    }
    base = atomicAdd(count, t);
  }
  __syncthreads();
  uint32_t o = base + part[threadIdx.x];
  for (uint32_t j = j0; j < j0 + E; ++j)
    if (j == 0 || k[j] != k[j - 1]) out[o++] = k[j];
}

// One workgroup: the sorted distinct keys of keys[0, n) (n <= kKeyTile) into
// table, their count into *K.
__global__ __launch_bounds__(kKeyThreads) void k_key_table(const unsigned long long *keys, uint32_t n,
                                                           unsigned long long *table, uint32_t *K) {
  __shared__ unsigned long long k[kKeyTile];
  __shared__ uint32_t part[kKeyThreads];
  load_key_tile(keys, n, 0, k);
  // thread t owns the contiguous positions [t E, t E + E), so the distinct
  // keys leave in sorted order
  constexpr uint32_t E = kKeyTile / kKeyThreads;
  const uint32_t j0 = threadIdx.x * E;
  uint32_t nd = 0;
  for (uint32_t j = j0; j < j0 + E; ++j) nd += (j == 0 || k[j] != k[j - 1]) ? 1u : 0u;
  part[threadIdx.x] = nd;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kKeyThreads; ++i) {
      const uint32_t v = part[i];
      part[i] = t;
      t += v;
    }
    *K = t;
  }
  __syncthreads();
  uint32_t o = part[threadIdx.x];
  for (uint32_t j = j0; j < j0 + E; ++j)
    if (j == 0 || k[j] != k[j - 1]) table[o++] = k[j];
}

// label[i] = index of keys[i] in the sorted table of K (<= kKeyTile) keys.
__global__ __launch_bounds__(kKeyThreads) void k_key_map(const unsigned long long *keys, uint32_t n,
                                                         const unsigned long long *table, const uint32_t *Kp,
                                                         uint32_t *label) {
  __shared__ unsigned long long t[kKeyTile];
  const uint32_t K = *Kp;
  for (uint32_t j = threadIdx.x; j < K; j += kKeyThreads) t[j] = table[j];
  __syncthreads();
  for (uint32_t i = blockIdx.x * kKeyThreads + threadIdx.x; i < n; i += gridDim.x * kKeyThreads) {
    const unsigned long long v = keys[i];
    uint32_t lo = 0, hi = K;  // first index with t[idx] >= v
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (t[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    label[i] = lo;
  }
}

// 32-bit keys widened to 64 bits.
__global__ void k_key_widen(const uint32_t *in, uint32_t n, unsigned long long *out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[i];
}

// ---- value code columns ------------------------------------------------------
// For feature f, the presorted row gives the distinct-value index d(s) of
// every sample (equal floats, +0 and -0 included, share an index).  When
// every pair of adjacent distinct values is further apart than the
// candidate gap ((double)a + 1e-5 < (double)b), a split position between
// two samples is a candidate exactly when their indices differ, and the
// split scan may read d(s) from a column of 1, 8 or 16 bits per sample
// instead of the 32-bit value.  Built from the presort's sorted keys and
// ids in three passes over tiles of kCodeTile positions: value-change flags
// and their in-tile prefix, a per-feature scan of the tile totals, and the
// scatter of d to sample order.  D[f] = distinct values, bad[f] = some
// adjacent pair closer than the gap.
constexpr int kCodeThreads = 256, kCodeItems = 4;
constexpr uint32_t kCodeTile = kCodeThreads * kCodeItems;

__device__ __forceinline__ float key_value(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Row-free features (two-valued, without rows): the threshold between the
// two values, (lowest + highest) * 0.5f without contraction, the value pair
// the split scan would meet at the one candidate position.
__global__ void k_rl_thr(const uint32_t *keys, uint32_t N, const int32_t *rl_feat, uint32_t Frl, float *thr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= Frl) return;
  const uint32_t *k = keys + (size_t)rl_feat[j] * N;
  thr[j] = __fmul_rn(__fadd_rn(key_value(k[0]), key_value(k[N - 1])), 0.5f);
}

// Per-sample signatures: bit j of sig[s] set when sample s holds the higher
// value of row-free feature j (its value-code bit).
__global__ void k_rl_sig(const uint8_t *codes, const uint64_t *fcode, const int32_t *rl_feat, uint32_t Frl,
                         uint32_t Wsig, uint32_t N, uint32_t *sig) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += gridDim.x * blockDim.x) {
    for (uint32_t w = 0; w < Wsig; ++w) {
      uint32_t word = 0;
      for (uint32_t b = 0; b < 32u && w * 32u + b < Frl; ++b) {
        const uint32_t *col = reinterpret_cast<const uint32_t *>(codes + fc_off(fcode[rl_feat[w * 32u + b]]));
        word |= ((col[s >> 5] >> (s & 31u)) & 1u) << b;
      }
      sig[(size_t)s * Wsig + w] = word;
    }
  }
}

__global__ __launch_bounds__(kCodeThreads) void k_code_flags(const uint32_t *keys, uint32_t N, uint32_t ntiles,
                                                              uint32_t *pre, uint32_t *tsum, uint32_t *bad) {
  __shared__ uint32_t ws[kCodeThreads / 64];
  const uint32_t f = blockIdx.y, tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const uint32_t *k = keys + (size_t)f * N;
  const uint32_t r0 = blockIdx.x * kCodeTile + tid * kCodeItems;
  uint32_t fl[kCodeItems], ts = 0, isbad = 0;
  uint32_t prev = (r0 > 0 && r0 <= N) ? k[r0 - 1] : 0u;
#pragma unroll
  for (int j = 0; j < kCodeItems; ++j) {
    const uint32_t r = r0 + j;
    fl[j] = 0;
    if (r < N) {
      const uint32_t cur = k[r];
      if (r > 0 && cur != prev) {
        fl[j] = 1;
        if (!((double)key_value(prev) + 0.00001 < (double)key_value(cur))) isbad = 1;
      }
      prev = cur;
    }
    ts += fl[j];
  }
  if (isbad) atomicOr(&bad[f], 1u);
  const uint32_t inc = wave_incl_scan(ts);
  if (lane == 63) ws[wave] = inc;
  __syncthreads();
  uint32_t run = inc - ts;
  for (uint32_t w = 0; w < wave; ++w) run += ws[w];
#pragma unroll
  for (int j = 0; j < kCodeItems; ++j) {
    run += fl[j];
    if (r0 + j < N) pre[(size_t)f * N + r0 + j] = run;  // inclusive, within the tile
  }
  if (tid == kCodeThreads - 1) tsum[(size_t)f * ntiles + blockIdx.x] = run;
}

// Exclusive scan of feature f's tile totals (one workgroup per feature);
// D[f] = total + 1.
__global__ __launch_bounds__(1024) void k_code_tscan(uint32_t *tsum, uint32_t ntiles, uint32_t *D) {
  __shared__ uint32_t ws[16];
  __shared__ uint32_t carry;
  const uint32_t f = blockIdx.x, tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  uint32_t *t = tsum + (size_t)f * ntiles;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < ntiles; b0 += 1024) {
    const uint32_t b = b0 + tid;
    const uint32_t v = b < ntiles ? t[b] : 0u;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    uint32_t off = carry + inc - v;
    for (uint32_t w = 0; w < wave; ++w) off += ws[w];
    if (b < ntiles) t[b] = off;
    __syncthreads();
    if (tid == 0) {
      uint32_t sum = 0;
      for (int w = 0; w < 16; ++w) sum += ws[w];
      carry += sum;
    }
    __syncthreads();
  }
  if (tid == 0) D[f] = carry + 1;
}

// d[f][id of position r] = in-tile prefix + the tile's offset.
__global__ __launch_bounds__(kCodeThreads) void k_code_scatter(const uint32_t *ids, const uint32_t *pre,
                                                                const uint32_t *tsum, uint32_t N, uint32_t ntiles,
                                                                uint32_t *d) {
  const uint32_t f = blockIdx.y;
  const uint32_t base = tsum[(size_t)f * ntiles + blockIdx.x];
  // each element stands alone here: consecutive threads take consecutive
  // positions (coalesced reads; the writes are a scatter by sample id)
#pragma unroll
  for (int j = 0; j < kCodeItems; ++j) {
    const uint32_t r = blockIdx.x * kCodeTile + (uint32_t)j * kCodeThreads + threadIdx.x;
    if (r < N) d[(size_t)f * N + ids[(size_t)f * N + r]] = base + pre[(size_t)f * N + r];
  }
}

// Packs d[f][.] into feature f's code column (kind 1: bits, 2: bytes, 3:
// 16-bit) at the byte offset fcode[f] names (kind 0: no column).
__global__ void k_code_pack(const uint32_t *d, uint32_t N, const uint64_t *fcode, uint8_t *codes) {
  const uint32_t f = blockIdx.y;
  const uint64_t fc = fcode[f];
  const uint32_t kind = fc_kind(fc);
  if (kind == 0) return;
  const uint32_t *src = d + (size_t)f * N;
  uint8_t *dst = codes + fc_off(fc);
  if (kind == 1) {
    // bits: one coalesced read per sample, a ballot per 64 samples, the
    // wave's two words written by lanes 0 and 32
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t *w32 = reinterpret_cast<uint32_t *>(dst);
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < N; base += gridDim.x * blockDim.x) {
      const uint32_t i = base + lane;
      const uint64_t m = ballot(i < N && (src[i] & 1u));
      if (lane == 0) w32[base / 32] = (uint32_t)m;
      if (lane == 32 && base + 32 < N) w32[base / 32 + 1] = (uint32_t)(m >> 32);
    }
    return;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    if (kind == 2) {
      dst[i] = (uint8_t)src[i];
    } else {
      reinterpret_cast<uint16_t *>(dst)[i] = (uint16_t)src[i];
    }
  }
}

// ---- blocked predict table ---------------------------------------------------------
// Built from the pre-order table (PNode, right-child indices global) and the
// node depths: flags of the block roots, their ranks by an exclusive scan,
// each node's slot and block root (the roots set their children, then the
// nodes one level below set theirs), then one scatter.

// flag[g] = 1 when node g roots a block (its depth is a multiple of 3).
__global__ void k_block_flags(const uint32_t *depth, uint32_t n, uint32_t *flag) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x)
    flag[g] = depth[g] % 3u == 0u ? 1u : 0u;
}

// Block roots: slot 0 and themselves as root; their children slots 1 and 2.
__global__ void k_block_top(const PNode *nodes, const uint32_t *depth, uint32_t n, uint8_t *slot, uint32_t *root) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    if (depth[g] % 3u != 0u) continue;
    slot[g] = 0;
    root[g] = g;
    const PNode p = nodes[g];
    if (p.feat < 0) continue;
    slot[g + 1] = 1;
    root[g + 1] = g;
    slot[p.right] = 2;
    root[p.right] = g;
  }
}

// Split nodes one level below a block root: their children's slots.
__global__ void k_block_mid(const PNode *nodes, const uint32_t *depth, uint32_t n, uint8_t *slot, uint32_t *root) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    if (depth[g] % 3u != 1u) continue;
    const PNode p = nodes[g];
    if (p.feat < 0) continue;
    const uint32_t s = slot[g], r = root[g];
    slot[g + 1] = (uint8_t)(2u * s + 1u);
    root[g + 1] = r;
    slot[p.right] = (uint8_t)(2u * s + 2u);
    root[p.right] = r;
  }
}

// Every node into its block's slot; split nodes at the bottom of a block
// point at their children's blocks.
__global__ void k_block_fill(const PNode *nodes, const uint32_t *depth, uint32_t n, const uint8_t *slot,
                             const uint32_t *root, const uint32_t *rank, BNode *blocks) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    const PNode p = nodes[g];
    BNode b;
    b.thr = p.thr;
    b.feat = p.feat;
    if (p.feat < 0) {
      b.a = p.value;
      b.b = 0u;
    } else if (depth[g] % 3u == 2u) {
      b.a = rank[g + 1];
      b.b = rank[p.right];
    } else {
      b.a = b.b = 0u;
    }
    blocks[(size_t)rank[root[g]] * kBlockSlots + slot[g]] = b;
  }
}

// Root block of every tree.
__global__ void k_block_roots(const uint32_t *first, uint32_t T, const uint32_t *rank, uint32_t *broots) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x)
    broots[t] = rank[first[t]];
}

// ---- predict ----------------------------------------------------------------
// Leaf label of sample x from node `node`: each level loads both children
// (the left one is the next node) with the split feature's value, so a
// level costs one memory latency rather than two.
// The blocked table: inside a block both children come from the line the
// block's first node fetched; at a block's bottom level both child blocks'
// first nodes are loaded with the split feature's value, so every level
// costs one memory latency and three levels one line from L2.
__device__ __forceinline__ uint32_t walk(const BNode *__restrict__ blocks, const float *__restrict__ x, uint32_t b) {
  const BNode *blk = blocks + (size_t)b * kBlockSlots;
  uint32_t s = 0;
  BNode nd = blk[0];
  while (nd.feat >= 0) {
    const float v = x[nd.feat];
    const bool inner = s < 3u;
    const BNode *lp = inner ? blk + 2u * s + 1u : blocks + (size_t)nd.a * kBlockSlots;
    const BNode *rp = inner ? blk + 2u * s + 2u : blocks + (size_t)nd.b * kBlockSlots;
    const BNode l = *lp, r = *rp;
    const bool left = v < nd.thr;
    if (inner) {
      s = left ? 2u * s + 1u : 2u * s + 2u;
    } else {
      blk = left ? lp : rp;
      s = 0;
    }
    nd = left ? l : r;
  }
  return nd.a;
}

__global__ __launch_bounds__(256) void k_predict(const BNode *__restrict__ nodes, const uint32_t *__restrict__ roots,
                                                 uint32_t T, const float *__restrict__ X, uint64_t R, uint32_t F,
                                                 uint32_t K, uint32_t *__restrict__ votes) {
  const uint64_t total = R * T;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = g % R;
    const uint32_t t = (uint32_t)(g / R);
    const float *x = X + r * F;
    atomicAdd(&votes[r * K + walk(nodes, x, roots[t])], 1u);
  }
}

// The same traversal with the trees split over the XCDs: workgroup b walks
// the trees t with t % kXcd == b % kXcd, so each XCD's L2 holds its eighth of
// the blocks.  Items (row, tree of the class) run tree-fastest, so the lanes
// of a wave share a row of X.  Needs gridDim.x a multiple of kXcd.
__global__ __launch_bounds__(256) void k_predict_xcd(const BNode *__restrict__ nodes,
                                                     const uint32_t *__restrict__ roots, uint32_t T,
                                                     const float *__restrict__ X, uint64_t R, uint32_t F, uint32_t K,
                                                     uint32_t *__restrict__ votes) {
  const uint32_t cls = blockIdx.x % kXcd;
  const uint32_t nt = (T - cls + kXcd - 1) / kXcd;  // trees of this class
  const uint64_t total = R * nt;
  const uint64_t step = (uint64_t)(gridDim.x / kXcd) * blockDim.x;
  const uint32_t lane = threadIdx.x % kWave;
  // whole waves step together, so that the lanes holding the same row and
  // label add their votes with one atomic
  for (uint64_t g0 = (uint64_t)(blockIdx.x / kXcd) * blockDim.x + (threadIdx.x - lane); g0 < total; g0 += step) {
    const uint64_t g = g0 + lane;
    const bool active = g < total;
    uint64_t key = 0;
    if (active) {
      const uint64_t r = g / nt;
      const uint32_t t = cls + kXcd * (uint32_t)(g % nt);
      key = r * K + walk(nodes, X + r * F, roots[t]);
    }
    uint64_t rem = ballot(active);
    while (rem) {
      const uint32_t leader = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
      const uint64_t kl = ((uint64_t)readlane((uint32_t)(key >> 32), leader) << 32) | readlane((uint32_t)key, leader);
      const uint64_t mm = ballot(active && key == kl);
      if (lane == leader) atomicAdd(&votes[kl], (uint32_t)__popcll(mm));
      rem &= ~mm;
    }
  }
}

// Winning label per row: most votes, lowest label on ties (np.argmax of
// np.bincount, random_forest.py:358).
__global__ void k_vote(const uint32_t *votes, uint64_t R, uint32_t K, uint32_t *labels) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *v = votes + r * K;
    uint32_t best = 0, bv = v[0];
    for (uint32_t k = 1; k < K; ++k)
      if (v[k] > bv) {
        bv = v[k];
        best = k;
      }
    labels[r] = best;
  }
}

}  // namespace cta

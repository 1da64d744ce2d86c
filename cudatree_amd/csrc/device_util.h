// device_util.h -- wave64 primitives, exact score comparison and the
// per-node feature permutation shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cta {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `m` below this lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  uint32_t lo = uniform((uint32_t)v), hi = uniform((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// ---- cross-lane moves by DPP (data-parallel primitives of the GFX9 VALU) --
// Each move is one VALU modifier instead of an LDS-crossbar permute.  A lane
// whose source lane is out of range, or whose row is masked off, reads 0.
// Controls: row_shr:n (0x110 + n, within rows of 16 lanes), row_bcast:15
// (0x142, lane 15 of each row to the next row), row_bcast:31 (0x143, lane 31
// to rows 2 and 3), wave_shl:1 (0x130, lane i reads lane i + 1).
constexpr int kDppRowShr = 0x110, kDppBcast15 = 0x142, kDppBcast31 = 0x143, kDppWaveShl1 = 0x130;

template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xF, true);
}

// Value of lane i + 1 (0 in lane 63).
__device__ __forceinline__ uint32_t wave_next(uint32_t v) { return dpp<kDppWaveShl1>(v); }

// Inclusive prefix sum over the 64 lanes: shifts by 1, 2, 4, 8 within each
// row of 16, then the row totals carried across rows.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp<kDppRowShr + 1>(v);
  v += dpp<kDppRowShr + 2>(v);
  v += dpp<kDppRowShr + 4>(v);
  v += dpp<kDppRowShr + 8>(v);
  v += dpp<kDppBcast15, 0xA>(v);
  v += dpp<kDppBcast31, 0xC>(v);
  return v;
}

// Reduction to lane 63 by the same steps; `comb` is associative and
// commutative, and lanes reading an out-of-range source see `zero`.
template <class T, class Move, class Comb>
__device__ __forceinline__ T wave_reduce_to_63(T v, Move mv, Comb comb) {
  v = comb(v, mv.template operator()<kDppRowShr + 1, 0xF>(v));
  v = comb(v, mv.template operator()<kDppRowShr + 2, 0xF>(v));
  v = comb(v, mv.template operator()<kDppRowShr + 4, 0xF>(v));
  v = comb(v, mv.template operator()<kDppRowShr + 8, 0xF>(v));
  v = comb(v, mv.template operator()<kDppBcast15, 0xA>(v));
  v = comb(v, mv.template operator()<kDppBcast31, 0xC>(v));
  return v;
}

struct Move64 {
  template <int kCtrl, int kRowMask>
  __device__ __forceinline__ uint64_t operator()(uint64_t v) const {
    return ((uint64_t)dpp<kCtrl, kRowMask>((uint32_t)(v >> 32)) << 32) | dpp<kCtrl, kRowMask>((uint32_t)v);
  }
};

__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), kWave - 1);
}

// Sum over the 64 lanes, returned in every lane.
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  v = wave_reduce_to_63(v, Move64{}, [](uint64_t x, uint64_t y) { return x + y; });
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, kWave - 1);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), kWave - 1);
  return ((uint64_t)hi << 32) | lo;
}

// ---- exact score comparison --------------------------------------------
// A split's score is Q = num/den with num = S_L n_R + S_R n_L and
// den = n_L n_R (S = sum of squared class counts).  Maximising Q is
// minimising the reference's weighted Gini (dfs_module.cu:404-407,
// bfs_module.cu:176-180).  num < 2^64 and den < 2^42 for n < 2^22, so the
// cross products need 128 bits.
__device__ __forceinline__ int q_cmp(uint64_t n1, uint64_t d1, uint64_t n2, uint64_t d2) {
  uint64_t ahi = __umul64hi(n1, d2), alo = n1 * d2;
  uint64_t bhi = __umul64hi(n2, d1), blo = n2 * d1;
  if (ahi != bhi) return ahi > bhi ? 1 : -1;
  if (alo != blo) return alo > blo ? 1 : -1;
  return 0;
}

// Candidate (score, feature, position); ok == 0 means "none".
struct Cand {
  uint64_t num, den;
  uint32_t pos;
  int32_t feat;
  uint32_t ok;
};

// True if a beats b: higher Q, then lower feature id, then lower position.
__device__ __forceinline__ bool cand_better(const Cand &a, const Cand &b) {
  if (!a.ok) return false;
  if (!b.ok) return true;
  int c = q_cmp(a.num, a.den, b.num, b.den);
  if (c) return c > 0;
  if (a.feat != b.feat) return a.feat < b.feat;
  return a.pos < b.pos;
}

struct MoveCand {
  template <int kCtrl, int kRowMask>
  __device__ __forceinline__ Cand operator()(const Cand &c) const {
    const Move64 m;
    Cand r;
    r.num = m.template operator()<kCtrl, kRowMask>(c.num);
    r.den = m.template operator()<kCtrl, kRowMask>(c.den);
    r.pos = dpp<kCtrl, kRowMask>(c.pos);
    r.feat = (int32_t)dpp<kCtrl, kRowMask>((uint32_t)c.feat);
    r.ok = dpp<kCtrl, kRowMask>(c.ok);  // out-of-range sources read ok = 0: never better
    return r;
  }
};

// Every lane ends with the best candidate of the wave (the order is total,
// so the reduction order does not matter).
__device__ __forceinline__ Cand wave_best(Cand c) {
  c = wave_reduce_to_63(c, MoveCand{}, [](const Cand &x, const Cand &y) { return cand_better(y, x) ? y : x; });
  Cand r;
  r.num = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(c.num >> 32), kWave - 1) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)c.num, kWave - 1);
  r.den = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(c.den >> 32), kWave - 1) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)c.den, kWave - 1);
  r.pos = (uint32_t)__builtin_amdgcn_readlane((int)c.pos, kWave - 1);
  r.feat = __builtin_amdgcn_readlane(c.feat, kWave - 1);
  r.ok = (uint32_t)__builtin_amdgcn_readlane((int)c.ok, kWave - 1);
  return r;
}

// ---- per-node feature permutation ---------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) keyed by the fit seed over the
// counter (start, depth, tree, "PERM") gives four round keys of a 4-round
// Feistel network on 2h bits (2^(2h) >= F); cycle walking restricts it to
// [0, F).  Visiting order of the features at the node whose segment starts
// at `start` in tree `tree` at depth `depth`.
struct Perm {
  uint32_t rk[4];
  uint32_t h, F;
  bool identity;
};

__device__ __forceinline__ void philox_round(uint32_t c[4], uint32_t k0, uint32_t k1) {
  uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
  uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
  uint32_t n0 = hi1 ^ c[1] ^ k0;
  uint32_t n2 = hi0 ^ c[3] ^ k1;
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    philox_round(c, k0, k1);
  }
}

__device__ __forceinline__ Perm make_perm(uint64_t seed, uint32_t tree, uint32_t depth, uint32_t start,
                                          uint32_t F, bool identity) {
  Perm p;
  p.F = F;
  p.identity = identity;
  uint32_t h = 1;
  while ((1ull << (2 * h)) < (uint64_t)F) ++h;
  p.h = h;
  uint32_t c[4] = {start, depth, tree, 0x5045524Du};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  p.rk[0] = c[0];
  p.rk[1] = c[1];
  p.rk[2] = c[2];
  p.rk[3] = c[3];
  return p;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint32_t feistel(uint32_t x, const Perm &p) {
  const uint32_t mask = (1u << p.h) - 1u;
  uint32_t L = x >> p.h, R = x & mask;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t f = fmix32(R ^ p.rk[r]) & mask;
    uint32_t nl = R;
    R = L ^ f;
    L = nl;
  }
  return (L << p.h) | R;
}

// Canonical bootstrap draws of tree `tree`: block `blk` gives draws
// 4*blk .. 4*blk+3 as the words of Philox4x32-10(counter = (blk, tree, "BOOT",
// 0), key = seed); draw u maps to sample floor(u * N / 2^32).
__device__ __forceinline__ void boot_block(uint64_t seed, uint32_t tree, uint32_t blk, uint32_t w[4]) {
  w[0] = blk;
  w[1] = tree;
  w[2] = 0x424F4F54u;
  w[3] = 0u;
  philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ __forceinline__ uint32_t draw_to_sample(uint32_t u, uint32_t N) {
  return (uint32_t)(((uint64_t)u * N) >> 32);
}

// pi(k) for k < F.
__device__ __forceinline__ uint32_t perm_at(const Perm &p, uint32_t k) {
  if (p.identity) return k;
  uint32_t y = feistel(k, p);
  while (y >= p.F) y = feistel(y, p);
  return y;
}

// Candidate test of the reference (common_func.cu:1, bfs_module.cu:173):
// float operands promoted to double against the double literal 1e-5.
__device__ __forceinline__ bool gap_ok(float a, float b) { return (double)a + 0.00001 < (double)b; }

}  // namespace cta

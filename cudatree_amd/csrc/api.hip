// api.hip -- host side of libcudatree_amd.so: the C-ABI declared in
// include/cudatree_amd.h, the fit driver (batched / hybrid / per-node
// schedules over the kernels of fit_kernels.hip), export to canonical
// pre-order trees, and prediction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cudatree_amd.h"
#include "aux_kernels.hip"
#include "fit_kernels.hip"
#include "subtree_kernels.hip"

using namespace cta;

// ---- errors ------------------------------------------------------------------
namespace {
thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(CTA_EHIP, "%s: %s", what, hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

template <class F>
int guarded(F &&f) {
  try {
    f();
    g_err.clear();
    return CTA_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return CTA_ENOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return CTA_EINTERNAL;
  }
}

// Device allocation owned by RAII.
struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  explicit DevMem(size_t n) { alloc(n); }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(CTA_ENOMEM, "hipMalloc of %zu bytes failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevMem() { release(); }
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

uint32_t grid_for(size_t n, int block = 256, uint32_t cap = 65535u * 16) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)std::min<size_t>(g, cap);
}
}  // namespace

// ---- context -------------------------------------------------------------------
struct KStat {
  uint64_t launches = 0;
  double ms = 0;
  uint64_t bytes = 0;
};

struct cta_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second fit lane: a context of its own (stream, workspaces, staging) on
  // the same device, created by the first fit that runs two lanes
  cta_ctx *lane = nullptr;
  void (*level_hook)(void *) = nullptr;  // called between groups of batched levels
  void *level_hook_arg = nullptr;
  std::recursive_mutex mu;
  int num_cu = 256;
  size_t lds_per_block = 65536;
  bool profiling = false;
  std::map<std::string, KStat> stats;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  // grow-only device workspaces and pinned host staging, reused across calls
  std::map<std::string, std::unique_ptr<DevMem>> wspace;
  std::map<std::string, std::pair<void *, size_t>> pinned;

  // Workspace bytes the context may hold (0: no limit); a growth beyond it
  // fails with CTA_ENOMEM instead of allocating.
  uint64_t ws_limit = 0;
  void ws_admit(size_t held, size_t want) const {
    if (ws_limit && ws_bytes() - held + want > ws_limit)
      fail(CTA_ENOMEM, "workspace of %zu more bytes would exceed the context's limit of %llu bytes", want - held,
           (unsigned long long)ws_limit);
  }
  void *ws(const std::string &key, size_t bytes) {
    auto &slot = wspace[key];
    if (!slot) slot.reset(new DevMem());
    if (slot->bytes < bytes || !slot->p) {
      ws_admit(slot->bytes, bytes);
      slot->alloc(bytes);
    }
    return slot->p;
  }
  template <class T>
  T *ws_as(const std::string &key, size_t count) {
    return static_cast<T *>(ws(key, count * sizeof(T)));
  }
  size_t ws_bytes() const {
    size_t b = 0;
    for (auto &kv : wspace) b += kv.second ? kv.second->bytes : 0;
    return b;
  }
  void ws_release() { wspace.clear(); }
  // Pinned host memory the device writes directly (mapped, coherent): the
  // host pointer; *dev receives the device's address for it.
  void *mapped(const std::string &key, size_t bytes, void **dev) {
    auto &slot = pinned[key];
    if (slot.second < bytes) {
      if (slot.first) (void)hipHostFree(slot.first);
      slot.first = nullptr;
      slot.second = 0;
      HIPC(hipHostMalloc(&slot.first, std::max<size_t>(bytes, 64), hipHostMallocMapped | hipHostMallocCoherent));
      slot.second = bytes;
    }
    HIPC(hipHostGetDevicePointer(dev, slot.first, 0));
    return slot.first;
  }
  void *host(const std::string &key, size_t bytes) {
    auto &slot = pinned[key];
    if (slot.second < bytes) {
      if (slot.first) (void)hipHostFree(slot.first);
      slot.first = nullptr;
      slot.second = 0;
      HIPC(hipHostMalloc(&slot.first, std::max<size_t>(bytes, 64), hipHostMallocDefault));
      slot.second = bytes;
    }
    return slot.first;
  }
  void host_time(const char *name, double ms) {
    if (!profiling) return;
    KStat &s = stats[name];
    s.launches += 1;
    s.ms += ms;
  }

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPC(hipEventCreate(&e));
    return e;
  }
  // Resolves recorded events (after a synchronisation).
  void resolve() {
    for (auto &p : pending) {
      float ms = 0;
      HIPC(hipEventSynchronize(p.b));
      HIPC(hipEventElapsedTime(&ms, p.a, p.b));
      KStat &s = stats[p.name];
      s.launches += 1;
      s.ms += ms;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void sync() {
    HIPC(hipStreamSynchronize(stream));
    if (profiling) resolve();
  }
  void add_bytes(const char *name, uint64_t b) {
    if (profiling) stats[name].bytes += b;
  }
};

// Brackets one launch with events on the context stream when profiling.
struct Timed {
  cta_ctx *c;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(cta_ctx *ctx, const char *n) : c(ctx), name(n) {
    if (c->profiling) {
      a = c->ev();
      b = c->ev();
      HIPC(hipEventRecord(a, c->stream));
    }
  }
  ~Timed() {
    if (c->profiling && a) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({name, a, b});
    }
  }
};

// ---- dataset -------------------------------------------------------------------
struct cta_dataset {
  cta_ctx *ctx = nullptr;
  int32_t N = 0, F = 0, K = 0, idx_bits = 32, lab_bits = 8;
  DevMem XT, y, si0;
  DevMem codes, fcode;  // value code columns (see k_code_index)
  // row-free features (two-valued, other than feature 0): tables of the fits
  // that keep no rows for them, and identity tables for fits that keep all
  int32_t Frl = 0, Wsig = 0, Fr = 0;
  DevMem rl_of, slot_of, row_feat, rl_thr, sig;
  DevMem ident, rl_none;
};

// ---- forest --------------------------------------------------------------------
struct cta_forest {
  cta_ctx *ctx = nullptr;
  int32_t T = 0, F = 0, K = 0;
  std::vector<int64_t> tree_off{0};
  std::vector<int32_t> feature;
  std::vector<float> threshold;
  std::vector<uint32_t> left, right, value, nsamp, depth;
  DevMem nodes, roots;     // pre-order predict table (PNode) and the trees' first nodes: the blocked table's source
  DevMem blocks, broots;   // blocked predict table (BNode) and each tree's root block
  bool device_ready = false;
};

// ---- allocation-free predict ---------------------------------------------------
// The vote matrix of cta_predict_device_noalloc lives in this file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it only names the
// symbol): each device's address comes from hipGetSymbolAddress, taken once
// per device when a context is created on it, and only that address is used.
static __device__ __align__(256) uint32_t scratch[CTA_NOALLOC_VOTE_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch[kMaxDevices] = {};

// Records this device's copy of `scratch` (device `device` is current).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;  // the no-alloc entry point rejects such a device
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(CTA_EHIP, "hipGetSymbolAddress(scratch) returned no address on device %d", device);
  g_scratch[device] = static_cast<uint32_t *>(p);
}

uint32_t *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}
}  // namespace

namespace {

cta_ctx *ctx_new(int device) {
  std::unique_ptr<cta_ctx> c(new cta_ctx());
  c->device = device;
  HIPC(hipSetDevice(device));
  HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  hipDeviceProp_t p;
  HIPC(hipGetDeviceProperties(&p, device));
  c->num_cu = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
  c->lds_per_block = p.sharedMemPerBlock > 0 ? p.sharedMemPerBlock : 65536;
  scratch_setup(device);
  return c.release();
}

// The context's second fit lane, created on first use.
cta_ctx *lane_ctx(cta_ctx *ctx) {
  if (!ctx->lane) ctx->lane = ctx_new(ctx->device);
  return ctx->lane;
}

// The blocked predict table of f from its pre-order table (f->nodes,
// f->roots, on the device) and the node depths (depth, device, pre-order);
// the pre-order table is released afterwards.  Temporaries come from the
// context's workspace.
void build_blocks(cta_ctx *ctx, cta_forest *f, const uint32_t *depth) {
  const uint32_t n = (uint32_t)f->tree_off.back(), T = (uint32_t)f->T;
  hipStream_t st = ctx->stream;
  uint32_t *flag = ctx->ws_as<uint32_t>("blk_flag", n), *rank = ctx->ws_as<uint32_t>("blk_rank", n);
  uint32_t *root = ctx->ws_as<uint32_t>("blk_root", n);
  uint8_t *slot = ctx->ws_as<uint8_t>("blk_slot", n);
  uint32_t *cnt = ctx->ws_as<uint32_t>("blk_cnt", 2);  // n, then the number of blocks
  uint32_t *bsum = ctx->ws_as<uint32_t>("blk_bsum", (size_t)n / kOffTile + 2);
  k_fill_u32<<<1, 64, 0, st>>>(cnt, 1, n);
  k_block_flags<<<grid_for(n), 256, 0, st>>>(depth, n, flag);
  // ranks of the block roots: an exclusive scan of the flags
  if (n <= 8 * kOffTile) {
    k_off_single<1><<<1, kOffSingleThreads, 0, st>>>(flag, cnt, 1u, rank, cnt + 1);
  } else {
    const uint32_t nb = (n + kOffTile - 1) / kOffTile;
    k_off_sums<1><<<nb, kOffThreads, 0, st>>>(flag, cnt, bsum);
    k_off_single<1><<<1, kOffSingleThreads, 0, st>>>(bsum, cnt, kOffTile, bsum, cnt + 1);
    k_off_apply<1><<<nb, kOffThreads, 0, st>>>(flag, cnt, bsum, rank);
  }
  const PNode *nodes = f->nodes.as<PNode>();
  k_block_top<<<grid_for(n), 256, 0, st>>>(nodes, depth, n, slot, root);
  k_block_mid<<<grid_for(n), 256, 0, st>>>(nodes, depth, n, slot, root);
  uint32_t nblocks = 0;
  HIPC(hipMemcpyAsync(&nblocks, cnt + 1, 4, hipMemcpyDeviceToHost, st));
  ctx->sync();
  f->blocks.alloc((size_t)std::max(nblocks, 1u) * kBlockSlots * sizeof(BNode));
  f->broots.alloc((size_t)std::max(T, 1u) * 4);
  k_block_fill<<<grid_for(n), 256, 0, st>>>(nodes, depth, n, slot, root, rank, f->blocks.as<BNode>());
  k_block_roots<<<grid_for(T), 256, 0, st>>>(f->roots.as<uint32_t>(), T, rank, f->broots.as<uint32_t>());
  HIPC(hipGetLastError());
  ctx->sync();
  f->nodes.release();
  f->roots.release();
}

void upload_forest(cta_forest *f) {
  if (f->device_ready) return;
  const size_t n = f->feature.size();
  std::vector<PNode> h(std::max<size_t>(n, 1));
  std::vector<uint32_t> depth(std::max<size_t>(n, 1), 0u);
  for (int32_t t = 0; t < f->T; ++t) {
    const int64_t base = f->tree_off[t];
    for (int64_t j = base; j < f->tree_off[t + 1]; ++j) {
      PNode &p = h[j];
      const bool leaf = f->left[j] == 0;
      p.thr = f->threshold[j];
      p.feat = leaf ? -1 : f->feature[j];
      p.right = leaf ? 0u : (uint32_t)(base + f->right[j]);
      p.value = f->value[j];
      // pre-order: children come after their parent
      if (!leaf) depth[j + 1] = depth[base + f->right[j]] = depth[j] + 1;
    }
  }
  std::vector<uint32_t> roots(std::max(f->T, 1));
  for (int32_t t = 0; t < f->T; ++t) roots[t] = (uint32_t)f->tree_off[t];
  f->nodes.alloc(sizeof(PNode) * h.size());
  f->roots.alloc(sizeof(uint32_t) * roots.size());
  cta_ctx *ctx = f->ctx;
  uint32_t *d_depth = ctx->ws_as<uint32_t>("blk_depth", depth.size());
  HIPC(hipMemcpyAsync(f->nodes.p, h.data(), sizeof(PNode) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(f->roots.p, roots.data(), sizeof(uint32_t) * roots.size(), hipMemcpyHostToDevice,
                      ctx->stream));
  HIPC(hipMemcpyAsync(d_depth, depth.data(), 4 * depth.size(), hipMemcpyHostToDevice, ctx->stream));
  build_blocks(ctx, f, d_depth);
  f->device_ready = true;
}

// ---- fit driver ------------------------------------------------------------------
struct HostEntry {
  FNode nd;
  std::vector<uint32_t> cnt;
};

// The reshuffle launch of a planned level (big and small tasks); the side
// bitmaps are staged in LDS unless `global_bits` or too large.
template <class IdxT, class LabT>
void launch_resh(cta_ctx *ctx, const FitArgs<IdxT, LabT> &a, bool global_bits, uint32_t grid,
                 const char *name = "k_resh") {
  const size_t lds = (((size_t)a.W + 3) & ~(size_t)3) * 4 + 2 * kSlots * 4;
  hipStream_t st = ctx->stream;
  Timed tm(ctx, name);
  if (!global_bits && lds <= ctx->lds_per_block) {
    static std::once_flag attr;
    std::call_once(attr, [&] {
      HIPC(hipFuncSetAttribute((const void *)k_resh<IdxT, LabT, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)ctx->lds_per_block));
    });
    k_resh<IdxT, LabT, true><<<grid, kReshThreads, lds, st>>>(a);
  } else {
    k_resh<IdxT, LabT, false><<<grid, kReshThreads, 2 * kSlots * 4, st>>>(a);
  }
}

// Device bytes one resident tree of a fit needs: ping-pong rows of the Fr
// features that keep rows, the side bitmap, and per sample
// pair the frontier, node arena and level-list entries.
size_t tree_bytes(int64_t N, int64_t Fr, int64_t K, int64_t m, size_t idx_bytes) {
  const size_t rows = 2 * (size_t)Fr * N * idx_bytes + N / 8 + 64;
  // (the level-list counts and offsets hold the larger of the two scans' fields)
  constexpr size_t kFields = kSelFields > kChFields ? kSelFields : kChFields;
  const size_t per_sample = 2 * sizeof(FNode) + 3 * 4 * (size_t)K + sizeof(NAux) + 4 * (size_t)m + 2 * 7 * 4 +
                            (sizeof(Task) + sizeof(ScanRes)) + 2 * sizeof(Task) + 2 * kFields * 4;
  return rows + per_sample * (size_t)N / 2 + 4096;
}

// Batched levels issued between two host reads of the control block.
#ifndef CTA_LEVEL_GROUP
#define CTA_LEVEL_GROUP 4
#endif
constexpr uint32_t kLevelGroup = CTA_LEVEL_GROUP;
// Default positions per split-scan task and elements per small-node
// reshuffle task, and positions per fill task (measured best at
// 581,012 x 54, 50 trees).
#ifndef CTA_FILL_CHUNK
#define CTA_FILL_CHUNK 512
#endif
constexpr uint32_t kScanChunk = 1024, kReshChunk = 4096, kFillChunk = CTA_FILL_CHUNK;


template <class IdxT, class LabT>
struct Fitter {
  cta_ctx *ctx;
  cta_dataset *ds;
  const cta_fit_opts &o;
  int32_t N, F, K, m;
  uint32_t Cs, Cf, Cr;
  uint32_t sub_max;  // small-subtree threshold of batched fits (0: none)
  hipStream_t st;
  uint32_t G;      // blocks of the wave-loop kernels
  uint32_t Gscan;  // blocks of k_scan: all resident at once, a multiple of kXcd
  bool scan_regs;  // k_scan keeps the class counts in lane registers
  uint64_t n_tot = 0;      // samples over the batch's trees
  bool fill_lds = false;   // side bitmaps built in LDS (N / 8 bytes fit a workgroup)
  // rows kept: every feature, or all but the row-free ones
  uint32_t Fr = 0, Frl = 0, Wsig = 0;
  const int32_t *rl_of_d = nullptr, *slot_of_d = nullptr, *row_feat_d = nullptr;
  uint64_t rl_cap = 0;  // frontier slots of each row-free count buffer
  uint32_t Jcap = 1;       // most fill workgroups per tree
  uint32_t *bsum = nullptr;

  Fitter(cta_ctx *c, cta_dataset *d, const cta_fit_opts &opts) : ctx(c), ds(d), o(opts) {
    N = d->N;
    F = d->F;
    K = d->K;
    m = opts.max_features;
    Cs = opts.scan_chunk > 0 ? (uint32_t)opts.scan_chunk : kScanChunk;
    Cr = opts.resh_chunk > 0 ? (uint32_t)opts.resh_chunk : kReshChunk;
    Cf = kFillChunk;
    sub_max = (opts.mode == CTA_MODE_BATCHED && opts.subtree_max > 0) ? (uint32_t)opts.subtree_max : 0u;
    st = c->stream;
#ifndef CTA_G_PER_CU
#define CTA_G_PER_CU 8
#endif
    G = (uint32_t)std::max(64, ctx->num_cu * CTA_G_PER_CU);
    const bool rowfree = !opts.all_rows && d->Frl > 0;
    Frl = rowfree ? (uint32_t)d->Frl : 0u;
    Wsig = rowfree ? (uint32_t)d->Wsig : 0u;
    Fr = rowfree ? (uint32_t)d->Fr : (uint32_t)F;
    rl_of_d = (rowfree ? d->rl_of : d->rl_none).as<int32_t>();
    slot_of_d = (rowfree ? d->slot_of : d->ident).as<int32_t>();
    row_feat_d = (rowfree ? d->row_feat : d->ident).as<int32_t>();
    // a wave-loop grid larger than what stays resident leaves a tail of
    // late workgroups; k_scan's registers allow fewer than 8 per CU
    int nb = 0;
    scan_regs = (uint32_t)K <= kScanRegsK && opts.hist_layout != CTA_HIST_LDS;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, scan_regs ? k_scan<IdxT, LabT, true> : k_scan<IdxT, LabT, false>, kBlock, smem()));
    nb = std::min(std::max(nb, 1), 8);
    Gscan = std::max<uint32_t>(kXcd, (uint32_t)(ctx->num_cu * nb) / kXcd * kXcd);
  }

  // side bitmap, then each wave's row-free counters
  size_t fill_lds_bytes() const { return (((size_t)N + 31) / 32) * 4 + (size_t)kFillWaves * Frl * K * 4; }

  size_t smem() const { return (size_t)kWavesPerBlock * 2 * K * sizeof(uint32_t); }

  // Bytes of device memory one resident tree needs (rows, marks, level state).
  size_t bytes_per_tree() const { return tree_bytes(N, Fr, K, m, sizeof(IdxT)); }

  // Grows trees [t0, t0 + Tb) and appends them to `out` in pre-order.
  void run_batch(int32_t t0, int32_t Tb, int32_t total_trees, const uint32_t *boot_ids, const int64_t *boot_off,
                 cta_forest *out) {
    const auto t_setup = std::chrono::steady_clock::now();
    const bool boot_gen = boot_ids == nullptr && o.bootstrap != 0;
    const bool boot = boot_ids != nullptr || boot_gen;
    const uint32_t W = ((uint32_t)N + 31u) / 32u;
    uint32_t *bits = ctx->ws_as<uint32_t>("bits", (size_t)Tb * W);
    std::vector<uint64_t> off(Tb + 1, 0);
    n_tot = 0;
    std::vector<uint32_t> gen_n;
    if (boot_gen) {
      // in-bag bitmaps drawn on the device; their counts size the batch
      uint32_t *cnts = ctx->ws_as<uint32_t>("boot_counts", Tb);
      k_fill_u32<<<grid_for(Tb), 256, 0, st>>>(cnts, Tb, 0u);
      if ((size_t)W * 4 <= (size_t)ctx->lds_per_block) {
        static std::once_flag attr;
        std::call_once(attr, [&] {
          HIPC(hipFuncSetAttribute((const void *)k_boot_gen_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)ctx->lds_per_block));
        });
        k_boot_gen_lds<<<Tb, 1024, (size_t)W * 4, st>>>(o.seed, (uint32_t)t0, N, W, bits);
      } else {
        k_fill_u32<<<grid_for((size_t)Tb * W), 256, 0, st>>>(bits, (size_t)Tb * W, 0u);
        k_boot_gen<<<dim3(grid_for((N + 3) / 4, 256, 256), Tb), 256, 0, st>>>(o.seed, (uint32_t)t0, N, W, bits);
      }
      k_bits_count<<<dim3(grid_for(W, 256, 8), Tb), 256, 0, st>>>(bits, W, cnts);  // few atomics per counter
      uint32_t *h = static_cast<uint32_t *>(ctx->host("boot_counts", (size_t)Tb * 4));
      HIPC(hipMemcpyAsync(h, cnts, (size_t)Tb * 4, hipMemcpyDeviceToHost, st));
      ctx->sync();
      gen_n.assign(h, h + Tb);
    }
    for (int32_t t = 0; t < Tb; ++t) {
      const uint64_t n = boot_gen ? gen_n[t]
                         : boot ? (uint64_t)(boot_off[t0 + t + 1] - boot_off[t0 + t])
                                : (uint64_t)N;
      off[t + 1] = off[t] + n;
      n_tot += n;
    }
    const uint64_t cap_q = n_tot / 2 + Tb + 64;
    const uint64_t cap_nodes = 2 * n_tot + Tb + 64;
    // split-scan tasks are sized per level
    const uint64_t cap_scan = units_for((uint64_t)Tb);
    const uint64_t cap_big = (uint64_t)Fr * (n_tot / kBigN + Tb + 1);
    const uint64_t cap_small = cap_q + 2ull * Fr * n_tot / Cr + 64;
    const uint64_t cap_fill = cap_q + n_tot / Cf + 64;
    const uint64_t cap_hist = (uint64_t)m * K * (2 * n_tot / Cs + 64);
    if (cap_scan > 0xFFFFFFF0ull || cap_small > 0xFFFFFFF0ull || cap_big > 0xFFFFFFF0ull ||
        cap_nodes > 0xFFFFFFF0ull || cap_hist > 0xFFFFFFF0ull)
      fail(CTA_EUNSUPPORTED, "batch of %d trees too large for 32-bit task indices", Tb);

    IdxT *buf0 = ctx->ws_as<IdxT>("buf0", (size_t)Tb * Fr * N), *buf1 = ctx->ws_as<IdxT>("buf1", (size_t)Tb * Fr * N);
    uint32_t *plan = ctx->ws_as<uint32_t>("plan", (size_t)Tb * 4);
    uint32_t *wg_off = ctx->ws_as<uint32_t>("wg_off", 2 * ((size_t)Tb + 1));
    uint32_t *tlo = ctx->ws_as<uint32_t>("tlo", (size_t)Tb * 2);
    constexpr uint64_t kFields = kSelFields > kChFields ? kSelFields : kChFields;
    uint32_t *vals = ctx->ws_as<uint32_t>("scan_vals", cap_q * kFields);
    uint32_t *offs = ctx->ws_as<uint32_t>("scan_offs", cap_q * kFields);
    bsum = ctx->ws_as<uint32_t>("scan_bsum", (cap_q / kOffTile + 2) * kFields);
    const size_t nbslot = kByteKinds * (size_t)kByteSlots * kByteStride;
    unsigned long long *bslot = ctx->ws_as<unsigned long long>("byte_slots", nbslot);
    fill_lds = !o.fill_global && fill_lds_bytes() <= (size_t)ctx->lds_per_block;
#ifndef CTA_FILL_JMUL
#define CTA_FILL_JMUL 1
#define CTA_FILL_JMAX 16
#endif
    Jcap = fill_lds ? (uint32_t)std::max<int64_t>(
                          1, std::min<int64_t>(CTA_FILL_JMAX, (CTA_FILL_JMUL * (int64_t)ctx->num_cu) / Tb))
                    : 1u;
    FNode *front[2] = {ctx->ws_as<FNode>("front0", cap_q), ctx->ws_as<FNode>("front1", cap_q)};
    uint32_t *cnt[2] = {ctx->ws_as<uint32_t>("cnt0", cap_q * K), ctx->ws_as<uint32_t>("cnt1", cap_q * K)};
    uint32_t *cntL = ctx->ws_as<uint32_t>("cntL", cap_q * K);
    NAux *aux = ctx->ws_as<NAux>("aux", cap_q);
    int32_t *feats = ctx->ws_as<int32_t>("feats", cap_q * m);
    Task *scan_tasks = ctx->ws_as<Task>("scan_tasks", cap_scan);
    ScanRes *res = ctx->ws_as<ScanRes>("res", cap_scan);
    uint32_t *hist = ctx->ws_as<uint32_t>("hist", cap_hist);
    Task *fill_tasks = ctx->ws_as<Task>("fill_tasks", cap_fill);
    Task *big_tasks = ctx->ws_as<Task>("big_tasks", cap_big);
    Task *small_tasks = ctx->ws_as<Task>("small_tasks", cap_small);
    int32_t *a_feat = ctx->ws_as<int32_t>("a_feat", cap_nodes);
    float *a_thr = ctx->ws_as<float>("a_thr", cap_nodes);
    uint32_t *a_left = ctx->ws_as<uint32_t>("a_left", cap_nodes), *a_right = ctx->ws_as<uint32_t>("a_right", cap_nodes);
    uint32_t *a_value = ctx->ws_as<uint32_t>("a_value", cap_nodes), *a_nsamp = ctx->ws_as<uint32_t>("a_nsamp", cap_nodes);
    uint32_t *a_depth = ctx->ws_as<uint32_t>("a_depth", cap_nodes);
    Ctl *ctl = ctx->ws_as<Ctl>("ctl", 1);
    uint32_t *root_cnt = ctx->ws_as<uint32_t>("root_cnt", (size_t)Tb * K);
    uint64_t *d_off = ctx->ws_as<uint64_t>("boot_off", Tb + 1);
    uint32_t *d_ids = nullptr;

    FitArgs<IdxT, LabT> a = {};
    a.X_T = ds->XT.as<float>();
    a.y = ds->y.as<LabT>();
    a.si0 = ds->si0.as<IdxT>();
    a.codes = ds->codes.as<uint8_t>();
    a.fcode = ds->fcode.as<uint64_t>();
    a.Frl = Frl;
    a.Wsig = Wsig;
    a.Fr = Fr;
    a.rl_of = rl_of_d;
    a.slot_of = slot_of_d;
    a.row_feat = row_feat_d;
    a.rl_thr = ds->rl_thr.as<float>();
    a.sig = ds->sig.as<uint32_t>();
    a.buf0 = buf0;
    a.buf1 = buf1;
    a.bits = bits;
    a.W = W;
    a.T = (uint32_t)Tb;
    a.N = N;
    a.F = F;
    a.K = K;
    a.m = m;
    a.mss = o.min_samples_split;
    a.max_depth = o.max_depth;
    a.identity = o.identity_perm;
    a.tree_base = (uint32_t)t0;
    a.seed = o.seed;
    a.Cs = Cs;
    a.Cf = Cf;
    a.Cr = Cr;
    a.hist_lds = o.hist_layout == CTA_HIST_LDS ? 1u : 0u;
    a.cntL = cntL;
    a.aux = aux;
    a.feats = feats;
    a.scan_tasks = scan_tasks;
    a.res = res;
    a.hist = hist;
    a.fill_tasks = fill_tasks;
    a.big_tasks = big_tasks;
    a.small_tasks = small_tasks;
    a.plan = plan;
    a.wg_off = wg_off;
    a.tlo = tlo;
    a.thi = tlo + Tb;
    a.vals = vals;
    a.offs = offs;
    a.bslot = bslot;
    a.arena = Arena{a_feat, a_thr, a_left, a_right, a_value, a_nsamp, a_depth};
    a.ctl = ctl;
    a.cap_q = (uint32_t)cap_q;
    a.cap_scan = (uint32_t)cap_scan;
    a.cap_hist = (uint32_t)cap_hist;
    a.cap_fill = (uint32_t)cap_fill;
    a.cap_big = (uint32_t)cap_big;
    a.cap_small = (uint32_t)cap_small;
    a.cap_nodes = (uint32_t)cap_nodes;
    // live children of at most sub_max samples: queued for k_subtree
    a.sub_max = sub_max;
    a.sub_queue = sub_max ? ctx->ws_as<FNode>("sub_queue", cap_q) : nullptr;
    a.cap_sub = sub_max ? (uint32_t)cap_q : 0u;

    // ---- setup: bootstrap compaction and roots
    Ctl hc = {};
    hc.n_nodes = (uint32_t)Tb;
    Ctl *hctl = static_cast<Ctl *>(ctx->host("ctl", sizeof(Ctl)));
    *hctl = hc;
    uint64_t *hoff = static_cast<uint64_t *>(ctx->host("boot_off", (Tb + 1) * 8));
    std::memcpy(hoff, off.data(), (Tb + 1) * 8);
    HIPC(hipMemcpyAsync(ctl, hctl, sizeof(Ctl), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d_off, hoff, (Tb + 1) * 8, hipMemcpyHostToDevice, st));
    k_fill_u32<<<grid_for((size_t)Tb * K), 256, 0, st>>>(root_cnt, (size_t)Tb * K, 0u);
    k_fill_u32<<<grid_for((size_t)Tb * 2), 256, 0, st>>>(tlo, (size_t)Tb * 2, 0u);
    k_fill_u32<<<grid_for(nbslot * 2), 256, 0, st>>>(reinterpret_cast<uint32_t *>(bslot), nbslot * 2, 0u);
    uint32_t src_root = 0;
    if (boot) {
      if (boot_gen) {
        k_root_counts_bits<LabT><<<dim3(grid_for(N, 256, 32), Tb), 256, K * 4, st>>>(a.y, bits, N, W, K, root_cnt);
      } else {
        uint32_t *ids = static_cast<uint32_t *>(ctx->host("boot_ids", n_tot * 4));
        for (int32_t t = 0; t < Tb; ++t)
          std::memcpy(ids + off[t], boot_ids + boot_off[t0 + t], (off[t + 1] - off[t]) * 4);
        d_ids = ctx->ws_as<uint32_t>("boot_ids", n_tot);
        HIPC(hipMemcpyAsync(d_ids, ids, n_tot * 4, hipMemcpyHostToDevice, st));
        k_fill_u32<<<grid_for((size_t)Tb * W), 256, 0, st>>>(bits, (size_t)Tb * W, 0u);
        k_boot_mark<<<dim3(grid_for(N, 256, 1024), Tb), 256, 0, st>>>(d_ids, d_off, Tb, W, bits);
        k_root_counts<LabT><<<dim3(grid_for(N, 256, 256), Tb), 256, K * 4, st>>>(a.y, d_ids, d_off, Tb, N, K,
                                                                                root_cnt);
      }
      a.front = front[0];
      k_ctl_set<<<1, 64, 0, st>>>(a.ctl, (uint32_t)Tb);
      k_boot_tasks<IdxT, LabT><<<grid_for((size_t)Tb * 64), 256, 0, st>>>(a, d_off, Tb, front[0]);
      offsets<kChFields>(a, (uint32_t)Tb, &a.ctl->c_ids);
      k_emit_resh<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
      reshuffle(a, true);
      src_root = 1;
    } else {
      k_root_counts<LabT><<<dim3(grid_for(N, 256, 256), Tb), 256, K * 4, st>>>(a.y, nullptr, nullptr, Tb, N, K,
                                                                              root_cnt);
    }
    a.next_front = front[1];
    a.next_cnt = cnt[1];
    k_roots<IdxT, LabT><<<1, kPlanThreads, 0, st>>>(a, root_cnt, boot ? d_off : nullptr, Tb, src_root);
    k_ctl_next<<<1, kByteKinds * kWave, 0, st>>>(a.ctl, a.bslot);
    int cur = 1;  // front[cur] holds the frontier
    rl_cap = 0;
    if (Frl) {
      uint64_t maxn = boot ? 0 : (uint64_t)N;
      for (int32_t t = 0; t < Tb && boot; ++t) maxn = std::max<uint64_t>(maxn, off[t + 1] - off[t]);
      ensure_rl(a, 2ull * Tb, cur);
      a.front = front[cur];
      rl_front(a, cur, (uint32_t)Tb, (uint32_t)maxn);
    }
    read_ctl(a, hc);
    if (hc.err) fail(CTA_EINTERNAL, "device capacity check failed during setup (err %u)", hc.err);
    if (boot) ctx->add_bytes("k_boot_compact", hc.bytes_boot);
    const auto t_levels = std::chrono::steady_clock::now();
    ctx->host_time("host_setup", std::chrono::duration<double, std::milli>(t_levels - t_setup).count());

    const int mode = o.mode;
    uint32_t depth_bound = 0;  // deepest node depth that can exist in the batch
    std::vector<uint32_t> level_off;  // batched mode: ids of depth d are [level_off[d], level_off[d+1])
    if (mode == CTA_MODE_BATCHED) {
      level_off = {0u, (uint32_t)Tb};
      uint32_t q = hc.q;
      const char *lvl_path = ctx->profiling ? std::getenv("CTA_LEVEL_LOG") : nullptr;
      FILE *lvl = lvl_path ? std::fopen(lvl_path, "a") : nullptr;
      if (lvl)
        std::fprintf(lvl, "level,q,scan_bytes,resh_big_bytes,resh_small_bytes,fill_bytes,k_scan_ms,k_resh_ms,"
                          "k_fill_ms,k_scan_hist_ms,k_select_ms,k_decide_ms,k_children_ms\n");
      int depth = 0;
      // Levels are issued in groups between host reads of the control block:
      // a frontier at most doubles per level, so the host sizes launches for
      // that bound, and levels issued after the frontier empties do nothing.
      const uint32_t group = lvl ? 1u : kLevelGroup;
      // The last level of a group copies the control block to host-mapped
      // memory and an event marks that copy, so the host reads the counts
      // and queues the next group while that level's reshuffle still runs.
      void *mirror_dev = nullptr;
      Ctl *mirror = lvl ? nullptr : static_cast<Ctl *>(ctx->mapped("ctl_mirror", sizeof(Ctl), &mirror_dev));
      struct EventOf {
        cta_ctx *c;
        hipEvent_t e;
        ~EventOf() { c->pool.push_back(e); }
      } ctl_ev{ctx, ctx->ev()};
      while (q > 0) {
        std::map<std::string, KStat> before;
        if (lvl) before = ctx->stats;
        const uint32_t d0 = hc.depth;
        uint32_t qb = q;
        for (uint32_t g = 0; g < group; ++g) {
          a.front = front[cur];
          a.cnt = cnt[cur];
          a.next_front = front[cur ^ 1];
          a.next_cnt = cnt[cur ^ 1];
          if (Frl) {
            ensure_rl(a, 2ull * qb, cur);
            a.rlc = rl_buf(cur);
            a.rlc_next = rl_buf(cur ^ 1);
            a.rl_small = rl_buf(2);
          }
          const bool last = g + 1 == group && mirror;
          a.ctl_host = last ? static_cast<Ctl *>(mirror_dev) : nullptr;
          level(a, qb, last ? ctl_ev.e : nullptr);
          a.ctl_host = nullptr;
          cur ^= 1;
          qb = (uint32_t)std::min<uint64_t>(2ull * qb, a.cap_q);
        }
        Ctl prev = hc;
        if (mirror) {
          HIPC(hipEventSynchronize(ctl_ev.e));
          std::memcpy(&hc, mirror, sizeof(Ctl));
        } else {
          read_ctl(a, hc);
        }
        account(prev, hc);
        if (hc.err) fail(CTA_EINTERNAL, "device capacity check failed (err %u)", hc.err);
        if (lvl) {
          auto d = [&](const char *k) { return ctx->stats[k].ms - before[k].ms; };
          std::fprintf(lvl, "%d,%u,%llu,%llu,%llu,%llu,%.4f,%.4f,%.4f,%.4f,%.4f,%.4f,%.4f\n", depth, q,
                       hc.bytes_scan - prev.bytes_scan, hc.bytes_resh - prev.bytes_resh,
                       hc.bytes_resh_small - prev.bytes_resh_small, hc.bytes_fill - prev.bytes_fill, d("k_scan"),
                       d("k_resh"), d("k_fill"), d("k_scan_hist"), d("k_select"), d("k_decide"), d("k_children"));
        }
        if (ctx->level_hook) ctx->level_hook(ctx->level_hook_arg);
        // levels of the group that had a frontier
        uint32_t qi = q;
        for (uint32_t d = d0 + 1; d <= hc.depth && qi > 0; ++d) {
          level_off.push_back(hc.lv_nodes[d % kLevelRing]);
          qi = hc.lv_q[d % kLevelRing];
          ++depth;
        }
        q = hc.q;
      }
      depth_bound = (uint32_t)depth;
      if (lvl) std::fclose(lvl);
      if (hc.n_sub) {
        subtrees(a, hc, Tb);
        depth_bound = std::max(depth_bound, hc.sub_depth);
        level_off.clear();  // subtree node ids are not grouped by depth
      }
    } else {
      // per-node (and hybrid) schedule: host stack of frontier entries
      std::vector<HostEntry> stack, queue;
      pull_frontier(front[cur], cnt[cur], hc.q, stack);
      std::reverse(stack.begin(), stack.end());
      const uint64_t thr = mode == CTA_MODE_HYBRID ? (uint64_t)std::max<int64_t>(o.bfs_threshold, 0) : 0;
      while (!stack.empty()) {
        HostEntry e = std::move(stack.back());
        stack.pop_back();
        if (mode == CTA_MODE_HYBRID && e.nd.n <= thr) {
          queue.push_back(std::move(e));
          continue;
        }
        depth_bound = std::max(depth_bound, e.nd.depth + 1);
        push_frontier(a, front[0], cnt[0], {e});
        a.front = front[0];
        a.cnt = cnt[0];
        a.next_front = front[1];
        a.next_cnt = cnt[1];
        if (Frl) {
          ensure_rl(a, 2, 0);
          rl_front(a, 0, 1, e.nd.n);
          a.rlc = rl_buf(0);
          a.rlc_next = rl_buf(1);
          a.rl_small = rl_buf(2);
        }
        level(a, 1);
        Ctl prev = hc;
        read_ctl(a, hc);
        account(prev, hc);
        if (hc.err) fail(CTA_EINTERNAL, "device capacity check failed (err %u)", hc.err);
        std::vector<HostEntry> kids;
        pull_frontier(front[1], cnt[1], hc.q, kids);
        for (auto it = kids.rbegin(); it != kids.rend(); ++it) stack.push_back(std::move(*it));
      }
      if (!queue.empty()) {
        uint32_t qd = 0;
        for (const auto &e : queue) qd = std::max(qd, e.nd.depth);
        // frontiers are kept sorted by tree
        std::stable_sort(queue.begin(), queue.end(),
                         [](const HostEntry &x, const HostEntry &y) { return x.nd.tree < y.nd.tree; });
        push_frontier(a, front[0], cnt[0], queue);
        cur = 0;
        uint32_t q = (uint32_t)queue.size();
        if (Frl) {
          uint32_t maxn = 0;
          for (const auto &e : queue) maxn = std::max(maxn, e.nd.n);
          ensure_rl(a, 2ull * q, cur);
          a.front = front[cur];
          rl_front(a, cur, q, maxn);
        }
        while (q > 0) {
          ++qd;
          a.front = front[cur];
          a.cnt = cnt[cur];
          a.next_front = front[cur ^ 1];
          a.next_cnt = cnt[cur ^ 1];
          if (Frl) {
            ensure_rl(a, 2ull * q, cur);
            a.rlc = rl_buf(cur);
            a.rlc_next = rl_buf(cur ^ 1);
            a.rl_small = rl_buf(2);
          }
          level(a, q);
          cur ^= 1;
          Ctl prev = hc;
          read_ctl(a, hc);
          account(prev, hc);
          if (hc.err) fail(CTA_EINTERNAL, "device capacity check failed (err %u)", hc.err);
          q = hc.q;
        }
        depth_bound = std::max(depth_bound, qd);
      }
    }
    const auto t_export = std::chrono::steady_clock::now();
    ctx->host_time("host_levels", std::chrono::duration<double, std::milli>(t_export - t_levels).count());
    export_batch(a, hc, depth_bound, level_off, Tb, Tb == total_trees, out);
    ctx->host_time("host_export",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_export).count());
  }

  // Every queued small node grown to its leaves by one k_subtree launch
  // (the side bitmaps start cleared; the waves keep them so).
  void subtrees(const FitArgs<IdxT, LabT> &a, Ctl &hc, int32_t Tb) {
    k_fill_u32<<<grid_for((size_t)Tb * a.W), 256, 0, st>>>(a.bits, (size_t)Tb * a.W, 0u);
    const size_t lds = (size_t)kSubWaves * sub_wave_words((uint32_t)K, Frl) * 4;
    int nb = 0;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_subtree<IdxT, LabT>, kSubThreads, lds));
    const uint32_t grid = (uint32_t)(ctx->num_cu * std::min(std::max(nb, 1), 8));
    {
      Timed tm(ctx, "k_subtree");
      k_subtree<IdxT, LabT><<<grid, kSubThreads, lds, st>>>(a);
    }
    HIPC(hipGetLastError());
    Ctl prev = hc;
    read_ctl(a, hc);
    account(prev, hc);
    if (hc.err) fail(CTA_EINTERNAL, "device capacity check failed in the small subtrees (err %u)", hc.err);
  }

  void read_ctl(const FitArgs<IdxT, LabT> &a, Ctl &hc) {
    Ctl *h = static_cast<Ctl *>(ctx->host("ctl", sizeof(Ctl)));
    HIPC(hipMemcpyAsync(h, a.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    ctx->sync();
    hc = *h;
  }

  void account(const Ctl &prev, const Ctl &now) {
    ctx->add_bytes("k_scan", (now.bytes_scan - prev.bytes_scan) - (now.bytes_scan_skip - prev.bytes_scan_skip));
    ctx->add_bytes("k_fill", now.bytes_fill - prev.bytes_fill);
    ctx->add_bytes("k_resh", (now.bytes_resh - prev.bytes_resh) + (now.bytes_resh_small - prev.bytes_resh_small));
    ctx->add_bytes("k_rl_count", now.bytes_rl - prev.bytes_rl);
    ctx->add_bytes("k_subtree", now.bytes_sub - prev.bytes_sub);
  }

  // Exclusive offsets and totals of the per-node counts in a.vals (NF
  // fields, ctl->q nodes) into a.offs and tot[0..NF); qb bounds ctl->q.
  template <int NF>
  void offsets(const FitArgs<IdxT, LabT> &a, uint32_t qb, uint32_t *tot) {
    const uint32_t *q = &a.ctl->q;
#ifndef CTA_OFF_SINGLE_TILES
#define CTA_OFF_SINGLE_TILES 2
#endif
    if (qb <= CTA_OFF_SINGLE_TILES * kOffTile) {
      k_off_single<NF><<<1, kOffSingleThreads, 0, st>>>(a.vals, q, 1u, a.offs, tot);
    } else {
      const uint32_t nb = (qb + kOffTile - 1) / kOffTile;
      k_off_sums<NF><<<nb, kOffThreads, 0, st>>>(a.vals, q, bsum);
      k_off_single<NF><<<1, kOffSingleThreads, 0, st>>>(bsum, q, kOffTile, bsum, tot);
      k_off_apply<NF><<<nb, kOffThreads, 0, st>>>(a.vals, q, bsum, a.offs);
    }
  }

  // Side bitmaps of the split nodes of a level of q frontier nodes.
  void fill(const FitArgs<IdxT, LabT> &a, uint32_t q) {
    Timed tm(ctx, "k_fill");
    if (fill_lds) {
      // fill tasks per tree (at most one per node plus one per Cf samples),
      // about kFillTasksPerGroup per workgroup, at most Jcap workgroups per tree
      const uint64_t est = ((uint64_t)q + n_tot / Cf) / std::max<uint32_t>(1, a.T);
      const uint32_t J = (uint32_t)std::min<uint64_t>(
          Jcap, std::max<uint64_t>(1, (est + kFillTasksPerGroup - 1) / kFillTasksPerGroup));
      static std::once_flag attr;
      std::call_once(attr, [&] {
        HIPC(hipFuncSetAttribute((const void *)k_fill_lds<IdxT, LabT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ctx->lds_per_block));
      });
      k_fill_lds<IdxT, LabT><<<a.T * J, kFillThreads, fill_lds_bytes(), st>>>(a, J);
    } else {
      k_fill_global<IdxT, LabT><<<G, kBlock, (size_t)kWavesPerBlock * Frl * K * 4, st>>>(a);
    }
  }

  // Row-free count buffers (frontier slots x Frl x K): one per frontier
  // parity and the split nodes' smaller-side counts (2); grown between
  // levels, the current one (`cur`) keeping its counts.
  static const char *rl_name(int which) { return which == 0 ? "rlc0" : which == 1 ? "rlc1" : "rlc2"; }
  uint32_t *rl_buf(int which) { return static_cast<uint32_t *>(ctx->wspace[rl_name(which)]->p); }
  void ensure_rl(const FitArgs<IdxT, LabT> &a, uint64_t need, int cur) {
    if (!Frl) return;
    need = std::min<uint64_t>(std::max<uint64_t>(need, 1), a.cap_q);
    if (need <= rl_cap) return;
    const size_t FK = (size_t)Frl * K;
    const uint64_t cap = std::min<uint64_t>(a.cap_q, std::max(need, 2 * rl_cap));
    ctx->sync();
    for (int w = 0; w < 3; ++w) {
      auto &slot = ctx->wspace[rl_name(w)];
      if (!slot) slot.reset(new DevMem());
      if (slot->p && slot->bytes >= cap * FK * 4) continue;
      ctx->ws_admit(slot->bytes, cap * FK * 4);
      DevMem fresh(cap * FK * 4);
      if (w == cur && rl_cap && slot->p)
        HIPC(hipMemcpyAsync(fresh.p, slot->p, rl_cap * FK * 4, hipMemcpyDeviceToDevice, st));
      ctx->sync();
      std::swap(slot->p, fresh.p);
      std::swap(slot->bytes, fresh.bytes);
    }
    rl_cap = cap;
  }
  // Counts of the frontier in a.front (q <= qb nodes of at most maxn
  // samples) into buffer `which`, from scratch.
  void rl_front(FitArgs<IdxT, LabT> a, int which, uint32_t qb, uint32_t maxn) {
    if (!Frl) return;
    const size_t FK = (size_t)Frl * K;
    uint32_t *buf = rl_buf(which);
    k_fill_u32<<<grid_for((size_t)qb * FK), 256, 0, st>>>(buf, (size_t)qb * FK, 0u);
    a.rlc = buf;
    // about 8 chunks per wave of the largest node
    const uint32_t per_block = Cf * kWavesPerBlock * 8;
    const uint32_t gx = std::min<uint32_t>(1024, std::max<uint32_t>(1, (maxn + per_block - 1) / per_block));
    Timed tm(ctx, "k_rl_count");
    k_rl_front<IdxT, LabT><<<dim3(gx, std::min<uint32_t>(qb, 65535u)), kBlock, (size_t)kWavesPerBlock * kRlCounters * 4,
                             st>>>(a);
  }

  // Split-scan tasks of a level of q nodes: nf * nch per multi-chunk node
  // (at most m n / Cs + m) and ceil(nf / fpt) <= m n / Cs + 1 per
  // single-chunk node, so at most q + 3 m n_tot / Cs; where fpt is held
  // down to give a small frontier kScanTaskTarget tasks, ceil(nf / fpt)
  // sums to at most q + 2 kScanTaskTarget.
  uint64_t units_for(uint64_t q) const {
    return q + 3 * (uint64_t)m * (n_tot / Cs + 1) + 2 * (uint64_t)kScanTaskTarget + 64;
  }

  // Grows the task and result lists for q nodes.
  void ensure_units(FitArgs<IdxT, LabT> &a, uint32_t q) {
    const uint64_t need = units_for(q);
    if (need <= a.cap_scan) return;
    const uint64_t cap = std::min<uint64_t>(0xFFFFFFF0ull, std::max(need, 2 * (uint64_t)a.cap_scan));
    if (need > cap) fail(CTA_EUNSUPPORTED, "too many split-scan units for 32-bit indices");
    ctx->sync();
    a.scan_tasks = ctx->ws_as<Task>("scan_tasks", cap);
    a.res = ctx->ws_as<ScanRes>("res", cap);
    a.cap_scan = (uint32_t)cap;
  }

  // One level over the frontier in a.front (ctl->q nodes, sorted by tree;
  // the host knows only a bound q on that count).
  // ev (optional): recorded after the level's plan kernel, which ends the
  // level (and copies the control block to a.ctl_host when set).
  void level(FitArgs<IdxT, LabT> &a, uint32_t q, hipEvent_t ev = nullptr) {
    const size_t sm = smem();
    ensure_units(a, q);
    {
      Timed tm(ctx, "k_select");
      k_select<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
      offsets<kSelFields>(a, q, &a.ctl->n_scan);
      k_select_emit<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
    }
    {
      Timed tm(ctx, "k_scan_hist");
      k_scan_hist<IdxT, LabT><<<G, kBlock, sm, st>>>(a);
    }
    k_scan_carry<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
    {
      Timed tm(ctx, "k_scan");
      if (scan_regs) k_scan<IdxT, LabT, true><<<Gscan, kBlock, sm, st>>>(a);
      else k_scan<IdxT, LabT, false><<<Gscan, kBlock, sm, st>>>(a);
    }
    {
      Timed tm(ctx, "k_decide");
      k_decide<IdxT, LabT><<<G, kBlock, sm, st>>>(a);
    }
    fill(a, q);
    {
      Timed tm(ctx, "k_children");
      offsets<kChFields>(a, q, &a.ctl->c_ids);
      k_children_emit<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
    }
    reshuffle(a, false, ev);  // its plan kernel also ends the level
    HIPC(hipGetLastError());
  }

  // Plan and partition of every row of the split nodes (task lists written);
  // `boot`: the bootstrap compaction, timed under its own name.
  void reshuffle(const FitArgs<IdxT, LabT> &a, bool boot = false, hipEvent_t ev = nullptr) {
    // the plan splits the work into about 4 units per CU; as many workgroups
    // (two stay resident per CU) let the CUs that finish early take more
#ifndef CTA_RESH_PER_CU
#define CTA_RESH_PER_CU 4
#endif
    const uint32_t grid = (uint32_t)std::max(8, ctx->num_cu * CTA_RESH_PER_CU);
    k_plan<IdxT, LabT><<<1, kPlanThreads, 0, st>>>(a, grid, boot ? 0u : 1u);
    if (ev) HIPC(hipEventRecord(ev, st));
    if (boot) launch_resh(ctx, a, o.resh_global != 0, grid, "k_boot_compact");
    else launch_resh(ctx, a, o.resh_global != 0, grid);
  }

  void pull_frontier(const FNode *fr, const uint32_t *c, uint32_t q, std::vector<HostEntry> &out) {
    if (q == 0) return;
    std::vector<FNode> nodes(q);
    std::vector<uint32_t> counts((size_t)q * K);
    HIPC(hipMemcpyAsync(nodes.data(), fr, q * sizeof(FNode), hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(counts.data(), c, (size_t)q * K * 4, hipMemcpyDeviceToHost, st));
    ctx->sync();
    for (uint32_t i = 0; i < q; ++i) {
      HostEntry e;
      e.nd = nodes[i];
      e.cnt.assign(counts.begin() + (size_t)i * K, counts.begin() + (size_t)(i + 1) * K);
      out.push_back(std::move(e));
    }
  }

  void push_frontier(const FitArgs<IdxT, LabT> &a, FNode *fr, uint32_t *c, const std::vector<HostEntry> &es) {
    std::vector<FNode> nodes;
    std::vector<uint32_t> counts;
    for (const auto &e : es) {
      nodes.push_back(e.nd);
      counts.insert(counts.end(), e.cnt.begin(), e.cnt.end());
    }
    HIPC(hipMemcpyAsync(fr, nodes.data(), nodes.size() * sizeof(FNode), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(c, counts.data(), counts.size() * 4, hipMemcpyHostToDevice, st));
    k_ctl_set<<<1, 64, 0, st>>>(a.ctl, (uint32_t)es.size());
    // host vectors must outlive the asynchronous copies
    ctx->sync();
  }

  // Canonical pre-order trees of the batch, numbered on the device, appended
  // to `out`; a single-batch fit also keeps the device predict table.
  // level_off: when node ids were allocated one depth at a time (batched
  // mode), depth d holds ids [level_off[d], level_off[d+1]) and no bucketing
  // is needed.
  void export_batch(const FitArgs<IdxT, LabT> &a, const Ctl &hc, uint32_t D, const std::vector<uint32_t> &level_off,
                    int32_t Tb, bool only_batch, cta_forest *out) {
    const uint32_t n = hc.n_nodes;
    const Arena &ar = a.arena;
    const bool contiguous = !level_off.empty();
    uint32_t *order = contiguous ? nullptr : ctx->ws_as<uint32_t>("canon_order", n);
    uint32_t *size = ctx->ws_as<uint32_t>("canon_size", n);
    uint32_t *pos = ctx->ws_as<uint32_t>("canon_pos", n);
    uint32_t *tr = ctx->ws_as<uint32_t>("canon_tr", n);
    uint32_t *toff = ctx->ws_as<uint32_t>("canon_toff", (size_t)Tb + 1);
    CanonOut co;
    co.feat = ctx->ws_as<int32_t>("canon_feat", n);
    co.thr = ctx->ws_as<float>("canon_thr", n);
    co.left = ctx->ws_as<uint32_t>("canon_left", n);
    co.right = ctx->ws_as<uint32_t>("canon_right", n);
    co.value = ctx->ws_as<uint32_t>("canon_value", n);
    co.nsamp = ctx->ws_as<uint32_t>("canon_nsamp", n);
    co.depth = ctx->ws_as<uint32_t>("canon_depth", n);
    co.pnode = ctx->ws_as<PNode>("canon_pnode", n);
    std::vector<uint32_t> doff;
    if (contiguous) {
      doff = level_off;
      D = (uint32_t)doff.size() - 2;
    } else {
      uint32_t *dcnt = ctx->ws_as<uint32_t>("canon_dcnt", D + 2);
      uint32_t *cursor = ctx->ws_as<uint32_t>("canon_cursor", D + 2);
      k_fill_u32<<<grid_for(D + 2), 256, 0, st>>>(dcnt, D + 2, 0u);
      k_depth_hist<<<grid_for(n), 256, 0, st>>>(ar.depth, n, dcnt);
      k_prefix_serial<<<1, 64, 0, st>>>(dcnt, D + 1);
      HIPC(hipMemcpyAsync(cursor, dcnt, (D + 2) * 4, hipMemcpyDeviceToDevice, st));
      k_depth_scatter<<<grid_for(n), 256, 0, st>>>(ar.depth, n, cursor, order);
      uint32_t *hoff = static_cast<uint32_t *>(ctx->host("canon_dcnt", (D + 2) * 4));
      HIPC(hipMemcpyAsync(hoff, dcnt, (D + 2) * 4, hipMemcpyDeviceToHost, st));
      ctx->sync();
      doff.assign(hoff, hoff + D + 2);
    }
    if (contiguous && D <= kLayoutMaxDepth) {
      // one workgroup per tree over all depths
      k_tree_layout<<<Tb, kLayoutThreads, 2 * ((size_t)D + 1) * 4, st>>>(ar, D, size, pos, tr);
      k_root_sizes<<<grid_for(Tb), 256, 0, st>>>(size, (uint32_t)Tb, toff);
      k_prefix_serial<<<1, 64, 0, st>>>(toff, (uint32_t)Tb);
    } else {
      for (int64_t d = D; d >= 0; --d) {
        const uint32_t b = doff[d], e = doff[d + 1];
        if (e > b) k_subtree_size<<<grid_for(e - b), 256, 0, st>>>(order, b, e, ar, size);
      }
      k_root_sizes<<<grid_for(Tb), 256, 0, st>>>(size, (uint32_t)Tb, toff);
      k_prefix_serial<<<1, 64, 0, st>>>(toff, (uint32_t)Tb);
      k_root_pos<<<grid_for(Tb), 256, 0, st>>>((uint32_t)Tb, pos, tr);
      for (uint32_t d = 0; d <= D; ++d) {
        const uint32_t b = doff[d], e = doff[d + 1];
        if (e > b) k_preorder<<<grid_for(e - b), 256, 0, st>>>(order, b, e, ar, size, pos, tr);
      }
    }
    k_canon_scatter<<<grid_for(n), 256, 0, st>>>(n, ar, size, pos, tr, toff, co);
    HIPC(hipGetLastError());
    // copy out through pinned staging: each array is appended to the forest
    // once its own copy has landed (an event per copy), the seven appends on
    // threads of their own, so they overlap the copies and each other
    uint32_t *stage = static_cast<uint32_t *>(ctx->host("export", (size_t)n * 4 * 7 + ((size_t)Tb + 1) * 4));
    uint32_t *htoff = stage + (size_t)7 * n;
    HIPC(hipMemcpyAsync(htoff, toff, ((size_t)Tb + 1) * 4, hipMemcpyDeviceToHost, st));
    const void *srcs[7] = {co.feat, co.thr, co.left, co.right, co.value, co.nsamp, co.depth};
    hipEvent_t evs[7];
    for (int k = 0; k < 7; ++k) {
      HIPC(hipMemcpyAsync(stage + (size_t)k * n, srcs[k], (size_t)n * 4, hipMemcpyDeviceToHost, st));
      evs[k] = ctx->ev();
      HIPC(hipEventRecord(evs[k], st));
    }
    if (only_batch) {
      out->nodes.alloc((size_t)n * sizeof(PNode));
      HIPC(hipMemcpyAsync(out->nodes.p, co.pnode, (size_t)n * sizeof(PNode), hipMemcpyDeviceToDevice, st));
      out->roots.alloc(((size_t)Tb + 1) * 4);
      HIPC(hipMemcpyAsync(out->roots.p, toff, (size_t)Tb * 4, hipMemcpyDeviceToDevice, st));
    }
    struct Done {  // returns the events to the context's pool
      cta_ctx *c;
      hipEvent_t *e;
      ~Done() {
        for (int k = 0; k < 7; ++k) c->pool.push_back(e[k]);
      }
    } done{ctx, evs};
    HIPC(hipEventSynchronize(evs[0]));  // the tree offsets landed before the first array
    if (htoff[Tb] != n) fail(CTA_EINTERNAL, "canonical export counted %u of %u nodes", htoff[Tb], n);
    std::atomic<bool> ok{true};
    auto app = [&](auto &vec, int k) {
      using V = typename std::decay_t<decltype(vec)>::value_type;
      if (hipEventSynchronize(evs[k]) != hipSuccess) {
        ok = false;
        return;
      }
      const V *p = reinterpret_cast<const V *>(stage + (size_t)k * n);
      try {
        vec.insert(vec.end(), p, p + n);
      } catch (...) {
        ok = false;
      }
    };
    {
      const std::function<void()> jobs[6] = {[&] { app(out->threshold, 1); }, [&] { app(out->left, 2); },
                                             [&] { app(out->right, 3); },     [&] { app(out->value, 4); },
                                             [&] { app(out->nsamp, 5); },     [&] { app(out->depth, 6); }};
      std::vector<std::thread> th;
      for (const auto &job : jobs) {
        try {
          th.emplace_back(job);
        } catch (...) {
          job();  // no thread to spare: this one here
        }
      }
      app(out->feature, 0);
      for (auto &t : th) t.join();
    }
    ctx->sync();
    if (!ok) fail(CTA_EINTERNAL, "copying the trees to host memory failed");
    const int64_t base = out->tree_off.back();
    for (int32_t t = 1; t <= Tb; ++t) out->tree_off.push_back(base + htoff[t]);
    out->T += Tb;
    if (only_batch) build_blocks(ctx, out, co.depth);
    out->device_ready = only_batch;
  }
};

// Trees each of two fit lanes needs at least for the lanes to pay.
constexpr int32_t kLaneMinTrees = 8;

// Appends trees [t, t + n) of `src` to `dst` (host arrays).
void append_trees(cta_forest *dst, const cta_forest *src, int32_t t, int32_t n) {
  const int64_t a = src->tree_off[t], b = src->tree_off[t + n];
  auto app = [&](auto &d, const auto &s) { d.insert(d.end(), s.begin() + a, s.begin() + b); };
  app(dst->feature, src->feature);
  app(dst->threshold, src->threshold);
  app(dst->left, src->left);
  app(dst->right, src->right);
  app(dst->value, src->value);
  app(dst->nsamp, src->nsamp);
  app(dst->depth, src->depth);
  const int64_t base = dst->tree_off.back();
  for (int32_t k = 1; k <= n; ++k) dst->tree_off.push_back(base + src->tree_off[t + k] - a);
  dst->T += n;
}

template <class IdxT, class LabT>
void fit_impl(cta_ctx *ctx, cta_dataset *ds, int32_t T, const uint32_t *ids, const int64_t *off,
              const cta_fit_opts &o, cta_forest *out) {
  Fitter<IdxT, LabT> fz(ctx, ds, o);
  // trees resident at once on one stream: as asked, else what 60% of the
  // free memory holds (workspaces are reused by this fit)
  int32_t per = o.trees_per_batch;
  if (o.mode != CTA_MODE_BATCHED) per = 1;
  if (per <= 0 && ctx->ws_limit) {
    // a context with a workspace limit: the trees the rest of it holds
    const size_t held = ctx->ws_bytes();
    const size_t avail = ctx->ws_limit > held ? (size_t)(ctx->ws_limit - held) : 0;
    per = (int32_t)std::max<size_t>(1, avail / fz.bytes_per_tree());
  } else if (per <= 0) {
    size_t fr = 0, tot = 0;
    HIPC(hipMemGetInfo(&fr, &tot));
    fr += ctx->ws_bytes() + (ctx->lane ? ctx->lane->ws_bytes() : 0);
    per = (int32_t)std::max<size_t>(1, (size_t)(0.6 * (double)fr) / fz.bytes_per_tree());
  }
  per = std::min(per, T);
  // two lanes: batches alternate between this context's stream and a second
  // one, so that one lane's short late levels overlap the other's work; each
  // holds half the trees (two batches resident at once), at least
  // kLaneMinTrees of them
  int32_t lanes = 1;
  if (o.mode == CTA_MODE_BATCHED && o.lanes == 2) {
    const int32_t half = std::min(o.trees_per_batch > 0 ? per : std::max<int32_t>(1, per / 2), (T + 1) / 2);
    if (half >= kLaneMinTrees && half < T) {
      lanes = 2;
      per = half;
    }
  }
  if (lanes == 1) {
    for (int32_t t0 = 0; t0 < T; t0 += per) fz.run_batch(t0, std::min(per, T - t0), T, ids, off, out);
    return;
  }
  std::vector<int32_t> b0;  // batch k: trees [b0[k], b0[k] + size), on lane k % 2
  for (int32_t t0 = 0; t0 < T; t0 += per) b0.push_back(t0);
  auto bsize = [&](size_t k) { return std::min(per, T - b0[k]); };
  cta_ctx *lc = lane_ctx(ctx);
  lc->profiling = ctx->profiling;
  lc->level_hook = ctx->level_hook;
  lc->level_hook_arg = ctx->level_hook_arg;
  cta_forest part[2];
  for (auto &p : part) {
    p.ctx = ctx;
    p.F = out->F;
    p.K = out->K;
  }
  std::exception_ptr lane_err;
  std::thread th([&] {
    try {
      HIPC(hipSetDevice(lc->device));
      Fitter<IdxT, LabT> fz1(lc, ds, o);
      for (size_t k = 1; k < b0.size(); k += 2) fz1.run_batch(b0[k], bsize(k), T, ids, off, &part[1]);
    } catch (...) {
      lane_err = std::current_exception();
    }
  });
  try {
    for (size_t k = 0; k < b0.size(); k += 2) fz.run_batch(b0[k], bsize(k), T, ids, off, &part[0]);
  } catch (...) {
    th.join();
    throw;
  }
  th.join();
  if (lane_err) std::rethrow_exception(lane_err);
  // the lane's kernel times join this context's (its host-side wall times
  // overlap this lane's and are left out)
  for (auto &kv : lc->stats) {
    if (kv.first.rfind("host_", 0) == 0) continue;
    KStat &d = ctx->stats[kv.first];
    d.launches += kv.second.launches;
    d.ms += kv.second.ms;
    d.bytes += kv.second.bytes;
  }
  lc->stats.clear();
  int32_t pos[2] = {0, 0};
  for (size_t k = 0; k < b0.size(); ++k) {
    append_trees(out, &part[k % 2], pos[k % 2], bsize(k));
    pos[k % 2] += bsize(k);
  }
  if (b0.size() == 2) {
    // one batch per lane: the predict table from the lanes' device tables
    // (the second's right-child indices shifted past the first's nodes)
    const int64_t n0 = part[0].tree_off.back(), n1 = part[1].tree_off.back();
    out->nodes.alloc(sizeof(PNode) * (size_t)std::max<int64_t>(n0 + n1, 1));
    HIPC(hipMemcpyAsync(out->nodes.p, ctx->ws_as<PNode>("canon_pnode", (size_t)n0), sizeof(PNode) * n0,
                        hipMemcpyDeviceToDevice, ctx->stream));
    if (n1)
      k_pnode_shift<<<grid_for((size_t)n1), 256, 0, ctx->stream>>>(lc->ws_as<PNode>("canon_pnode", (size_t)n1),
                                                                    (uint32_t)n1, (uint32_t)n0,
                                                                    out->nodes.as<PNode>() + n0);
    uint32_t *hr = static_cast<uint32_t *>(ctx->host("lane_roots", ((size_t)T + 1) * 4));
    for (int32_t t = 0; t <= T; ++t) hr[t] = (uint32_t)out->tree_off[t];
    out->roots.alloc(((size_t)T + 1) * 4);
    HIPC(hipMemcpyAsync(out->roots.p, hr, ((size_t)T + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    // the node depths of both parts, in the same order
    uint32_t *dep = ctx->ws_as<uint32_t>("blk_depth", (size_t)std::max<int64_t>(n0 + n1, 1));
    HIPC(hipMemcpyAsync(dep, ctx->ws_as<uint32_t>("canon_depth", (size_t)n0), 4 * (size_t)n0, hipMemcpyDeviceToDevice,
                        ctx->stream));
    if (n1) {
      // the second lane's copy lands on its own stream first
      lc->sync();
      HIPC(hipMemcpyAsync(dep + n0, lc->ws_as<uint32_t>("canon_depth", (size_t)n1), 4 * (size_t)n1,
                          hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPC(hipGetLastError());
    build_blocks(ctx, out, dep);
    out->device_ready = true;
  }
}

void build_codes(cta_ctx *ctx, cta_dataset *ds, const uint32_t *keys, const uint32_t *ids);

template <class IdxT>
void presort(cta_ctx *ctx, cta_dataset *ds) {
  const uint32_t N = ds->N, F = ds->F;
  const size_t total = (size_t)N * F;
  hipStream_t st = ctx->stream;
  uint32_t *k0 = ctx->ws_as<uint32_t>("sort_k0", total), *k1 = ctx->ws_as<uint32_t>("sort_k1", total);
  uint32_t *v0 = ctx->ws_as<uint32_t>("sort_v0", total), *v1 = ctx->ws_as<uint32_t>("sort_v1", total);
  const uint32_t ntiles = (N + kSortTile - 1) / kSortTile;
  uint32_t *hist = ctx->ws_as<uint32_t>("sort_hist", (size_t)F * kDigits * ntiles);
  // the digits that vary over each feature's keys: a pass over a digit that
  // all keys share keeps their order, so a feature takes only the others.
  // A feature with two distinct keys takes one pass, over the digit holding
  // the highest bit where they differ: that digit orders them (the bits
  // above it agree), so the pass is the stable partition.
  uint32_t *mask = ctx->ws_as<uint32_t>("sort_mask", 5 * (size_t)F);
  k_fill_u32<<<grid_for(F), 256, 0, st>>>(mask, F, 0u);
  k_fill_u32<<<grid_for(2 * (size_t)F), 256, 0, st>>>(mask + F, 2 * (size_t)F, ~0u);
  k_fill_u32<<<grid_for(2 * (size_t)F), 256, 0, st>>>(mask + 3 * F, 2 * (size_t)F, 0u);
  const dim3 rgrid(std::min<uint32_t>(64, (N + 255) / 256), F);
  k_key_mask<<<rgrid, 256, 0, st>>>(ds->XT.as<float>(), N, F, mask);
  k_key_others<<<rgrid, 256, 0, st>>>(ds->XT.as<float>(), N, F, mask);
  constexpr uint32_t kPasses = 32 / 4;
  // host table: the 5 F reductions, then odd[F], then the pass lists
  // (kPasses x F)
  uint32_t *h = static_cast<uint32_t *>(ctx->host("sort_plan", ((size_t)5 * F + F + kPasses * F) * 4));
  HIPC(hipMemcpyAsync(h, mask, 5 * (size_t)F * 4, hipMemcpyDeviceToHost, st));
  ctx->sync();
  uint32_t *odd = h + 5 * F, *lists = odd + F;
  uint32_t nlist[kPasses] = {};
  std::vector<uint32_t> in1(F), passes(F);  // passes: bit s set = pass s taken
  for (uint32_t f = 0; f < F; ++f) {
    const uint32_t vary = h[f] ^ h[F + f], lo = h[2 * F + f], hi = h[3 * F + f], others = h[4 * F + f];
    uint32_t p = 0;
    if (lo != hi && others == 0) {
      p = 1u << ((31u - (uint32_t)__builtin_clz(lo ^ hi)) / 4u);
    } else {
      for (uint32_t s = 0; s < kPasses; ++s) p |= ((vary >> (4 * s)) & 15u) ? 1u << s : 0u;
    }
    passes[f] = p;
    odd[f] = (uint32_t)__builtin_popcount(p) & 1u;
    in1[f] = odd[f];  // where the keys start
  }
  for (uint32_t s = 0; s < kPasses; ++s)
    for (uint32_t f = 0; f < F; ++f)
      if ((passes[f] >> s) & 1u) {
        lists[(size_t)s * F + nlist[s]++] = f | (in1[f] << 31);
        in1[f] ^= 1u;
      }
  uint32_t *dplan = ctx->ws_as<uint32_t>("sort_plan", (size_t)F + kPasses * F);
  HIPC(hipMemcpyAsync(dplan, odd, ((size_t)F + kPasses * F) * 4, hipMemcpyHostToDevice, st));
  k_sort_init<<<grid_for(total), 256, 0, st>>>(ds->XT.as<float>(), N, F, dplan, k0, v0, k1, v1);
  {
    Timed tm(ctx, "k_presort");
    for (uint32_t s = 0; s < kPasses; ++s) {
      if (!nlist[s]) continue;
      const SortPass sp = {dplan + F + (size_t)s * F, k0, v0, k1, v1};
      k_sort_hist<<<dim3(ntiles, nlist[s]), kSortThreads, 0, st>>>(sp, N, ntiles, 4 * s, hist);
      k_sort_scan<<<nlist[s], kSortThreads, 0, st>>>(sp, hist, ntiles);
      k_sort_scatter<<<dim3(ntiles, nlist[s]), kSortThreads, 0, st>>>(sp, N, ntiles, 4 * s, hist);
    }
  }
  // every feature's sorted keys and ids are in buffer 0
  const uint32_t *kin = k0, *vin = v0;
  ds->si0.alloc(total * sizeof(IdxT));
  if (ds->lab_bits == 8)
    k_sort_finish<IdxT, uint8_t><<<grid_for(total), 256, 0, st>>>(vin, total, ds->y.as<uint8_t>(), ds->si0.as<IdxT>());
  else
    k_sort_finish<IdxT, uint16_t><<<grid_for(total), 256, 0, st>>>(vin, total, ds->y.as<uint16_t>(),
                                                                   ds->si0.as<IdxT>());
  build_codes(ctx, ds, kin, vin);
  HIPC(hipGetLastError());
  ctx->sync();
}

// Value code columns: per feature, the width of its distinct-value index
// (1, 8 or 16 bits) when adjacent distinct values are all further apart than
// the candidate gap, else none (the split scan reads X_T).  keys / ids: the
// presort's sorted keys and sample ids, F x N.
void build_codes(cta_ctx *ctx, cta_dataset *ds, const uint32_t *keys, const uint32_t *ids) {
  const uint32_t N = ds->N, F = ds->F;
  hipStream_t st = ctx->stream;
  const uint32_t ntiles = (N + kCodeTile - 1) / kCodeTile;
  uint32_t *pre = ctx->ws_as<uint32_t>("code_pre", (size_t)N * F);
  uint32_t *d = ctx->ws_as<uint32_t>("code_index", (size_t)N * F);
  uint32_t *tsum = ctx->ws_as<uint32_t>("code_tsum", (size_t)ntiles * F);
  uint32_t *db = ctx->ws_as<uint32_t>("code_meta", 2 * (size_t)F);
  k_fill_u32<<<grid_for(2 * (size_t)F), 256, 0, st>>>(db, 2 * (size_t)F, 0u);
  k_code_flags<<<dim3(ntiles, F), kCodeThreads, 0, st>>>(keys, N, ntiles, pre, tsum, db + F);
  k_code_tscan<<<F, 1024, 0, st>>>(tsum, ntiles, db);
  k_code_scatter<<<dim3(ntiles, F), kCodeThreads, 0, st>>>(ids, pre, tsum, N, ntiles, d);
  uint32_t *h = static_cast<uint32_t *>(ctx->host("code_meta", 2 * (size_t)F * 4));
  HIPC(hipMemcpyAsync(h, db, 2 * (size_t)F * 4, hipMemcpyDeviceToHost, st));
  ctx->sync();
  std::vector<uint64_t> fc(F);
  uint64_t off = 0;
  for (uint32_t f = 0; f < F; ++f) {
    const uint32_t D = h[f], bad = h[F + f];
    uint32_t kind = 0;
    uint64_t bytes = 0;
    if (!bad && D <= 2) {
      kind = 1;
      bytes = ((uint64_t)N + 31) / 32 * 4;
    } else if (!bad && D <= 256) {
      kind = 2;
      bytes = N;
    } else if (!bad && D <= 65536) {
      kind = 3;
      bytes = 2 * (uint64_t)N;
    }
    fc[f] = off | kind;
    off += (bytes + 255) & ~(uint64_t)255;
  }
  ds->codes.alloc(std::max<uint64_t>(off, 256));
  ds->fcode.alloc((size_t)F * 8);
  uint64_t *hf = static_cast<uint64_t *>(ctx->host("code_fcode", (size_t)F * 8));
  std::memcpy(hf, fc.data(), (size_t)F * 8);
  HIPC(hipMemcpyAsync(ds->fcode.p, hf, (size_t)F * 8, hipMemcpyHostToDevice, st));
  k_code_pack<<<dim3(grid_for(N, 256, 256), F), 256, 0, st>>>(d, N, ds->fcode.as<uint64_t>(), ds->codes.as<uint8_t>());
  // Row-free features: the two-valued ones other than feature 0 (which keeps
  // its row: leaf labels and the children's samples are read from it), as
  // many as the count kernels' per-wave counters hold (2 x Frl x K).
  const uint32_t K = ds->K;
  const uint32_t max_rl = kRlCounters / (2 * K);
  std::vector<int32_t> tab((size_t)5 * F + 1, -1);  // rl_of, slot_of, ident, none, then row_feat / rl_feat
  int32_t *rl_of = tab.data(), *slot_of = rl_of + F, *ident = slot_of + F;  // then none: all -1
  std::vector<int32_t> row_feat, rl_feat;
  for (uint32_t f = 0; f < F; ++f) {
    ident[f] = (int32_t)f;
    if (f != 0 && fc_kind(fc[f]) == 1 && rl_feat.size() < max_rl) {
      rl_of[f] = (int32_t)rl_feat.size();
      rl_feat.push_back((int32_t)f);
    } else {
      slot_of[f] = (int32_t)row_feat.size();
      row_feat.push_back((int32_t)f);
    }
  }
  ds->Frl = (int32_t)rl_feat.size();
  ds->Wsig = (ds->Frl + 31) / 32;
  ds->Fr = (int32_t)row_feat.size();
  int32_t *ht = static_cast<int32_t *>(ctx->host("rl_tables", ((size_t)6 * F + 2) * 4));
  std::memcpy(ht, tab.data(), (size_t)4 * F * 4);
  std::memcpy(ht + 4 * F, row_feat.data(), row_feat.size() * 4);
  std::memcpy(ht + 5 * F, rl_feat.data(), rl_feat.size() * 4);
  auto upload = [&](DevMem &m, const int32_t *h, size_t n) {
    m.alloc(std::max<size_t>(n, 1) * 4);
    if (n) HIPC(hipMemcpyAsync(m.p, h, n * 4, hipMemcpyHostToDevice, st));
  };
  upload(ds->rl_of, ht, F);
  upload(ds->slot_of, ht + F, F);
  upload(ds->ident, ht + 2 * F, F);
  upload(ds->rl_none, ht + 3 * F, F);
  upload(ds->row_feat, ht + 4 * F, row_feat.size());
  DevMem rlf;
  upload(rlf, ht + 5 * F, rl_feat.size());
  ds->rl_thr.alloc(std::max<size_t>(ds->Frl, 1) * 4);
  ds->sig.alloc(std::max<size_t>((size_t)N * ds->Wsig, 1) * 4);
  if (ds->Frl) {
    k_rl_thr<<<grid_for(ds->Frl), 256, 0, st>>>(keys, N, rlf.as<int32_t>(), (uint32_t)ds->Frl, ds->rl_thr.as<float>());
    k_rl_sig<<<grid_for(N), 256, 0, st>>>(ds->codes.as<uint8_t>(), ds->fcode.as<uint64_t>(), rlf.as<int32_t>(),
                                          (uint32_t)ds->Frl, (uint32_t)ds->Wsig, N, ds->sig.as<uint32_t>());
  }
  HIPC(hipGetLastError());
  ctx->sync();
}

void dataset_build(cta_ctx *ctx, cta_dataset *ds, const float *X_dev, const uint32_t *y_dev) {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t N = ds->N, F = ds->F;
  hipStream_t st = ctx->stream;
  uint32_t *err = ctx->ws_as<uint32_t>("ds_err", 1);
  k_fill_u32<<<1, 64, 0, st>>>(err, 1, 0u);
  ds->XT.alloc((size_t)N * F * 4);
  k_transpose<<<dim3((N + kTT - 1) / kTT, (F + kTT - 1) / kTT), 256, 0, st>>>(X_dev, N, F, ds->XT.as<float>(), err);
  if (ds->lab_bits == 8) {
    ds->y.alloc(N);
    k_labels<uint8_t><<<grid_for(N), 256, 0, st>>>(y_dev, N, ds->K, ds->y.as<uint8_t>(), err);
  } else {
    ds->y.alloc((size_t)N * 2);
    k_labels<uint16_t><<<grid_for(N), 256, 0, st>>>(y_dev, N, ds->K, ds->y.as<uint16_t>(), err);
  }
  uint32_t *hp = static_cast<uint32_t *>(ctx->host("ds_err", 4));
  HIPC(hipMemcpyAsync(hp, err, 4, hipMemcpyDeviceToHost, st));
  ctx->sync();
  const uint32_t herr = *hp;
  if (herr & 1u) fail(CTA_EINVAL, "X contains NaN");
  if (herr & 2u) fail(CTA_EINVAL, "a label is outside [0, K)");
  if (ds->idx_bits == 16) presort<uint16_t>(ctx, ds);
  else presort<uint32_t>(ctx, ds);
  ctx->host_time("host_dataset",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

void check_sizes(int32_t N, int32_t F, int32_t K) {
  if (N <= 0 || F <= 0 || K <= 0) fail(CTA_EINVAL, "N, F and K must be positive (got %d, %d, %d)", N, F, K);
  if (N > CTA_MAX_SAMPLES) fail(CTA_EUNSUPPORTED, "N = %d exceeds %d samples", N, CTA_MAX_SAMPLES);
  if (F > CTA_MAX_FEATURES) fail(CTA_EUNSUPPORTED, "F = %d exceeds %d features", F, CTA_MAX_FEATURES);
  if (K > CTA_MAX_CLASSES) fail(CTA_EUNSUPPORTED, "K = %d exceeds %d classes", K, CTA_MAX_CLASSES);
}

}  // namespace

// ---- C-ABI -------------------------------------------------------------------------
extern "C" {

const char *cta_last_error(void) { return g_err.c_str(); }

int cta_version(void) { return 20000; }

int cta_ctx_create(int device, cta_ctx **out) {
  return guarded([&] {
    if (!out) fail(CTA_EINVAL, "out is NULL");
    int n = 0;
    HIPC(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(CTA_EINVAL, "device %d not present (%d devices)", device, n);
    *out = ctx_new(device);
  });
}

int cta_ctx_destroy(cta_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->lane) {
      cta_ctx_destroy(ctx->lane);
      ctx->lane = nullptr;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    ctx->wspace.clear();
    for (auto &kv : ctx->pinned)
      if (kv.second.first) (void)hipHostFree(kv.second.first);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int cta_ctx_synchronize(cta_ctx *ctx) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    ctx->sync();
  });
}

int cta_ctx_mem_info(cta_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    HIPC(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    HIPC(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  });
}

static int dataset_create_common(cta_ctx *ctx, const float *X, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                                 int32_t idx_bits, bool device, cta_dataset **out) {
  return guarded([&] {
    if (!ctx || !X || !y || !out) fail(CTA_EINVAL, "NULL argument");
    check_sizes(N, F, K);
    if (idx_bits == 0) idx_bits = N <= 65535 ? 16 : 32;
    if (idx_bits != 16 && idx_bits != 32) fail(CTA_EINVAL, "idx_bits must be 0, 16 or 32");
    if (idx_bits == 16 && N > 65535) fail(CTA_EINVAL, "16-bit indices need N <= 65535");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    std::unique_ptr<cta_dataset> ds(new cta_dataset());
    ds->ctx = ctx;
    ds->N = N;
    ds->F = F;
    ds->K = K;
    ds->idx_bits = idx_bits;
    ds->lab_bits = K <= 255 ? 8 : 16;
    if (device) {
      dataset_build(ctx, ds.get(), X, y);
    } else {
      DevMem dX((size_t)N * F * 4), dy((size_t)N * 4);
      HIPC(hipMemcpyAsync(dX.p, X, (size_t)N * F * 4, hipMemcpyHostToDevice, ctx->stream));
      HIPC(hipMemcpyAsync(dy.p, y, (size_t)N * 4, hipMemcpyHostToDevice, ctx->stream));
      dataset_build(ctx, ds.get(), dX.as<float>(), dy.as<uint32_t>());
    }
    *out = ds.release();
  });
}

// Distinct keys of the device keys k[0, N) in ascending order into table
// (device, kKeyTile entries); returns their count.  Rounds of
// k_key_distinct shrink the list until one tile holds it: a tile of at most
// CTA_MAX_CLASSES distinct keys leaves at most half of its kKeyTile, so a
// round that leaves more than n / 2 + kKeyTile keys has met more classes
// than the limit.
uint32_t key_classes(cta_ctx *ctx, const unsigned long long *k, uint32_t N, unsigned long long *table) {
  static_assert(kKeyTile >= 2 * CTA_MAX_CLASSES, "a key tile holds twice the class limit");
  hipStream_t st = ctx->stream;
  uint32_t *cnt = ctx->ws_as<uint32_t>("key_count", 2);
  uint32_t *hc = static_cast<uint32_t *>(ctx->host("key_count", 8));
  const unsigned long long *cur = k;
  unsigned long long *buf[2] = {nullptr, nullptr};
  uint32_t n = N;
  int w = 0;
  while (n > kKeyTile) {
    // a round writes at most one key per input key
    if (!buf[w]) buf[w] = ctx->ws_as<unsigned long long>(w ? "key_b" : "key_a", n);
    k_fill_u32<<<1, 64, 0, st>>>(cnt, 1, 0u);
    k_key_distinct<<<(n + kKeyTile - 1) / kKeyTile, kKeyThreads, 0, st>>>(cur, n, buf[w], cnt);
    HIPC(hipMemcpyAsync(hc, cnt, 4, hipMemcpyDeviceToHost, st));
    ctx->sync();
    const uint32_t n2 = hc[0];
    if (n2 > n / 2 + kKeyTile) fail(CTA_EUNSUPPORTED, "more than %d distinct class values", CTA_MAX_CLASSES);
    cur = buf[w];
    n = n2;
    w ^= 1;
  }
  k_key_table<<<1, kKeyThreads, 0, st>>>(cur, n, table, cnt + 1);
  HIPC(hipMemcpyAsync(hc + 1, cnt + 1, 4, hipMemcpyDeviceToHost, st));
  ctx->sync();
  if (hc[1] > (uint32_t)CTA_MAX_CLASSES) fail(CTA_EUNSUPPORTED, "more than %d distinct class values", CTA_MAX_CLASSES);
  return hc[1];
}

int cta_dataset_create_keys(cta_ctx *ctx, const float *X, const void *keys, int32_t key_bits, int32_t N, int32_t F,
                            int32_t idx_bits, int32_t *K_out, uint64_t *classes, cta_dataset **out) {
  return guarded([&] {
    if (!ctx || !X || !keys || !K_out || !classes || !out) fail(CTA_EINVAL, "NULL argument");
    if (key_bits != 32 && key_bits != 64) fail(CTA_EINVAL, "key_bits must be 32 or 64");
    check_sizes(N, F, 1);
    if (idx_bits == 0) idx_bits = N <= 65535 ? 16 : 32;
    if (idx_bits != 16 && idx_bits != 32) fail(CTA_EINVAL, "idx_bits must be 0, 16 or 32");
    if (idx_bits == 16 && N > 65535) fail(CTA_EINVAL, "16-bit indices need N <= 65535");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevMem dX((size_t)N * F * 4), dk((size_t)N * 8), dy((size_t)N * 4);
    HIPC(hipMemcpyAsync(dX.p, X, (size_t)N * F * 4, hipMemcpyHostToDevice, st));
    if (key_bits == 64) {
      HIPC(hipMemcpyAsync(dk.p, keys, (size_t)N * 8, hipMemcpyHostToDevice, st));
    } else {
      HIPC(hipMemcpyAsync(dy.p, keys, (size_t)N * 4, hipMemcpyHostToDevice, st));
      k_key_widen<<<grid_for(N), 256, 0, st>>>(dy.as<uint32_t>(), (uint32_t)N, dk.as<unsigned long long>());
    }
    unsigned long long *table = ctx->ws_as<unsigned long long>("key_table", kKeyTile);
    const uint32_t K = key_classes(ctx, dk.as<unsigned long long>(), (uint32_t)N, table);
    k_key_map<<<grid_for(N, kKeyThreads, 1024), kKeyThreads, 0, st>>>(dk.as<unsigned long long>(), (uint32_t)N,
                                                                       table, ctx->ws_as<uint32_t>("key_count", 2) + 1,
                                                                       dy.as<uint32_t>());
    HIPC(hipMemcpyAsync(classes, table, (size_t)K * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipGetLastError());
    ctx->sync();
    check_sizes(N, F, (int32_t)K);
    std::unique_ptr<cta_dataset> ds(new cta_dataset());
    ds->ctx = ctx;
    ds->N = N;
    ds->F = F;
    ds->K = (int32_t)K;
    ds->idx_bits = idx_bits;
    ds->lab_bits = K <= 255 ? 8 : 16;
    dataset_build(ctx, ds.get(), dX.as<float>(), dy.as<uint32_t>());
    *K_out = (int32_t)K;
    *out = ds.release();
  });
}

int cta_dataset_create(cta_ctx *ctx, const float *X, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                       int32_t idx_bits, cta_dataset **out) {
  return dataset_create_common(ctx, X, y, N, F, K, idx_bits, false, out);
}

int cta_dataset_create_device(cta_ctx *ctx, const float *X_dev, const uint32_t *y_dev, int32_t N, int32_t F,
                              int32_t K, int32_t idx_bits, cta_dataset **out) {
  return dataset_create_common(ctx, X_dev, y_dev, N, F, K, idx_bits, true, out);
}

int cta_dataset_destroy(cta_dataset *ds) {
  return guarded([&] {
    if (!ds) return;
    std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
    (void)hipSetDevice(ds->ctx->device);
    (void)hipStreamSynchronize(ds->ctx->stream);
    delete ds;
  });
}

int cta_dataset_info(const cta_dataset *ds, int32_t *N, int32_t *F, int32_t *K, int32_t *idx_bits) {
  return guarded([&] {
    if (!ds) fail(CTA_EINVAL, "ds is NULL");
    if (N) *N = ds->N;
    if (F) *F = ds->F;
    if (K) *K = ds->K;
    if (idx_bits) *idx_bits = ds->idx_bits;
  });
}

int cta_dataset_sorted(cta_dataset *ds, uint32_t *out) {
  return guarded([&] {
    if (!ds || !out) fail(CTA_EINVAL, "NULL argument");
    cta_ctx *ctx = ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    const size_t total = (size_t)ds->N * ds->F;
    DevMem w(total * 4);
    if (ds->idx_bits == 16)
      k_widen<uint16_t><<<grid_for(total), 256, 0, ctx->stream>>>(ds->si0.as<uint16_t>(), total, w.as<uint32_t>());
    else
      k_widen<uint32_t><<<grid_for(total), 256, 0, ctx->stream>>>(ds->si0.as<uint32_t>(), total, w.as<uint32_t>());
    HIPC(hipMemcpyAsync(out, w.p, total * 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

void cta_fit_opts_default(cta_fit_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->seed = 0;
  o->max_features = 0;
  o->min_samples_split = 1;
  o->max_depth = 0;
  o->identity_perm = 0;
  o->mode = CTA_MODE_BATCHED;
  o->bfs_threshold = 0;
}

int cta_fit(cta_ctx *ctx, cta_dataset *ds, int32_t T, const uint32_t *boot_ids, const int64_t *boot_off,
            const cta_fit_opts *opts, cta_forest **out) {
  return guarded([&] {
    if (!ctx || !ds || !opts || !out) fail(CTA_EINVAL, "NULL argument");
    if (ds->ctx != ctx) fail(CTA_EINVAL, "dataset belongs to another context");
    if (T <= 0) fail(CTA_EINVAL, "T must be positive");
    cta_fit_opts o = *opts;
    if (o.max_features <= 0 || o.max_features > ds->F)
      fail(CTA_EINVAL, "max_features must be in [1, %d] (got %d)", ds->F, o.max_features);
    if (o.mode < CTA_MODE_BATCHED || o.mode > CTA_MODE_PER_NODE) fail(CTA_EINVAL, "unknown mode %d", o.mode);
    if (o.scan_chunk < 0 || o.resh_chunk < 0) fail(CTA_EINVAL, "chunk sizes must be >= 0");
    if ((o.scan_chunk && o.scan_chunk < 64) || (o.resh_chunk && o.resh_chunk < 64))
      fail(CTA_EINVAL, "chunk sizes must be 0 or >= 64");
    if (o.hist_layout < CTA_HIST_AUTO || o.hist_layout > CTA_HIST_LDS)
      fail(CTA_EINVAL, "unknown hist_layout %d", o.hist_layout);
    if (o.hist_layout == CTA_HIST_REGS && ds->K > 64)
      fail(CTA_EINVAL, "hist_layout CTA_HIST_REGS needs K <= 64 (K = %d)", ds->K);
    if (boot_ids) {
      if (!boot_off) fail(CTA_EINVAL, "boot_off is NULL");
      for (int32_t t = 0; t < T; ++t) {
        const int64_t a = boot_off[t], b = boot_off[t + 1];
        if (b <= a || b - a > ds->N) fail(CTA_EINVAL, "tree %d has an empty or oversized in-bag set", t);
        for (int64_t j = a; j < b; ++j)
          if (boot_ids[j] >= (uint32_t)ds->N || (j > a && boot_ids[j] <= boot_ids[j - 1]))
            fail(CTA_EINVAL, "in-bag ids of tree %d are not sorted unique ids below N", t);
      }
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    std::unique_ptr<cta_forest> f(new cta_forest());
    f->ctx = ctx;
    f->F = ds->F;
    f->K = ds->K;
    const bool b16 = ds->idx_bits == 16, l8 = ds->lab_bits == 8;
    if (b16 && l8) fit_impl<uint16_t, uint8_t>(ctx, ds, T, boot_ids, boot_off, o, f.get());
    else if (b16) fit_impl<uint16_t, uint16_t>(ctx, ds, T, boot_ids, boot_off, o, f.get());
    else if (l8) fit_impl<uint32_t, uint8_t>(ctx, ds, T, boot_ids, boot_off, o, f.get());
    else fit_impl<uint32_t, uint16_t>(ctx, ds, T, boot_ids, boot_off, o, f.get());
    upload_forest(f.get());  // the predict table is resident when the fit returns
    *out = f.release();
  });
}

int cta_forest_destroy(cta_forest *f) {
  return guarded([&] {
    if (!f) return;
    std::lock_guard<std::recursive_mutex> lk(f->ctx->mu);
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    delete f;
  });
}

int cta_forest_info(const cta_forest *f, int32_t *n_trees, int64_t *n_nodes, int32_t *F, int32_t *K) {
  return guarded([&] {
    if (!f) fail(CTA_EINVAL, "forest is NULL");
    if (n_trees) *n_trees = f->T;
    if (n_nodes) *n_nodes = f->tree_off.back();
    if (F) *F = f->F;
    if (K) *K = f->K;
  });
}

int cta_forest_export(const cta_forest *f, int64_t *tree_off, int32_t *feature, float *threshold, uint32_t *left,
                      uint32_t *right, uint32_t *value, uint32_t *n_samples, uint32_t *depth) {
  return guarded([&] {
    if (!f || !tree_off) fail(CTA_EINVAL, "NULL argument");
    std::memcpy(tree_off, f->tree_off.data(), f->tree_off.size() * 8);
    const size_t n = f->feature.size();
    if (feature) std::memcpy(feature, f->feature.data(), n * 4);
    if (threshold) std::memcpy(threshold, f->threshold.data(), n * 4);
    if (left) std::memcpy(left, f->left.data(), n * 4);
    if (right) std::memcpy(right, f->right.data(), n * 4);
    if (value) std::memcpy(value, f->value.data(), n * 4);
    if (n_samples) {
      if (f->nsamp.size() == n) std::memcpy(n_samples, f->nsamp.data(), n * 4);
      else std::memset(n_samples, 0, n * 4);
    }
    if (depth) {
      if (f->depth.size() == n) std::memcpy(depth, f->depth.data(), n * 4);
      else std::memset(depth, 0, n * 4);
    }
  });
}

// Checks exported arrays (pre-order layout, ids and labels in range) and
// makes a resident forest of them; n_samples / depth may be NULL.
static cta_forest *import_arrays(cta_ctx *ctx, int32_t T, int32_t F, int32_t K, const int64_t *tree_off,
                                 const int32_t *feature, const float *threshold, const uint32_t *left,
                                 const uint32_t *right, const uint32_t *value, const uint32_t *nsamp,
                                 const uint32_t *depth) {
  if (T <= 0 || F <= 0 || K <= 0 || F > CTA_MAX_FEATURES || K > CTA_MAX_CLASSES) fail(CTA_EINVAL, "bad sizes");
  if (tree_off[0] != 0) fail(CTA_EINVAL, "tree_off[0] must be 0");
  for (int32_t t = 0; t < T; ++t) {
    const int64_t a = tree_off[t], b = tree_off[t + 1];
    if (b <= a) fail(CTA_EINVAL, "tree %d is empty", t);
    const int64_t n = b - a;
    for (int64_t j = a; j < b; ++j) {
      const bool leaf = left[j] == 0 || right[j] == 0;
      if (leaf) {
        if (value[j] >= (uint32_t)K) fail(CTA_EINVAL, "leaf label %u out of range", value[j]);
      } else {
        if (feature[j] < 0 || feature[j] >= F) fail(CTA_EINVAL, "feature %d out of range", feature[j]);
        // pre-order: left child next, right child later, both inside the tree
        if ((int64_t)left[j] != j - a + 1 || (int64_t)right[j] <= j - a + 1 || (int64_t)right[j] >= n)
          fail(CTA_EINVAL, "tree %d node %lld is not in pre-order layout", t, (long long)(j - a));
      }
    }
  }
  std::unique_ptr<cta_forest> f(new cta_forest());
  f->ctx = ctx;
  f->T = T;
  f->F = F;
  f->K = K;
  const int64_t n = tree_off[T];
  f->tree_off.assign(tree_off, tree_off + T + 1);
  f->feature.assign(feature, feature + n);
  f->threshold.assign(threshold, threshold + n);
  f->left.assign(left, left + n);
  f->right.assign(right, right + n);
  f->value.assign(value, value + n);
  if (nsamp) f->nsamp.assign(nsamp, nsamp + n);
  if (depth) f->depth.assign(depth, depth + n);
  for (int64_t j = 0; j < n; ++j)
    if (f->left[j] == 0 || f->right[j] == 0) f->left[j] = f->right[j] = 0;
  upload_forest(f.get());
  return f.release();
}

int cta_forest_import(cta_ctx *ctx, int32_t T, int32_t F, int32_t K, const int64_t *tree_off, const int32_t *feature,
                      const float *threshold, const uint32_t *left, const uint32_t *right, const uint32_t *value,
                      cta_forest **out) {
  return guarded([&] {
    if (!ctx || !tree_off || !feature || !threshold || !left || !right || !value || !out)
      fail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    *out = import_arrays(ctx, T, F, K, tree_off, feature, threshold, left, right, value, nullptr, nullptr);
  });
}

// ---- serialised forests ------------------------------------------------------------
// Byte image (little-endian): a 48-byte header {magic "CTAFORST", u32
// version, u32 header bytes, i32 T, F, K, u32 flags, i64 nodes, u64 0}, then
// tree_off (T + 1 x i64) and the node arrays feature (i32), threshold (f32),
// left, right, value, n_samples, depth (u32), n entries each.
namespace {
constexpr char kForestMagic[8] = {'C', 'T', 'A', 'F', 'O', 'R', 'S', 'T'};
constexpr uint32_t kForestHeader = 48;
struct ForestHeader {
  char magic[8];
  uint32_t version, header_bytes;
  int32_t T, F, K;
  uint32_t flags;
  int64_t nodes;
  uint64_t reserved;
};
static_assert(sizeof(ForestHeader) == kForestHeader, "forest header layout");
uint64_t forest_image_bytes(int64_t T, int64_t n) { return kForestHeader + 8 * (uint64_t)(T + 1) + 28 * (uint64_t)n; }
}  // namespace

int cta_forest_save(const cta_forest *f, void *buf, uint64_t cap, uint64_t *size) {
  return guarded([&] {
    if (!f || !size) fail(CTA_EINVAL, "NULL argument");
    const int64_t n = f->tree_off.back();
    const uint64_t need = forest_image_bytes(f->T, n);
    *size = need;
    if (!buf) return;
    if (cap < need) fail(CTA_EINVAL, "buffer of %llu bytes holds less than the %llu-byte image",
                         (unsigned long long)cap, (unsigned long long)need);
    ForestHeader h = {};
    std::memcpy(h.magic, kForestMagic, 8);
    h.version = CTA_FOREST_FORMAT;
    h.header_bytes = kForestHeader;
    h.T = f->T;
    h.F = f->F;
    h.K = f->K;
    h.nodes = n;
    char *p = static_cast<char *>(buf);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    std::memcpy(p, f->tree_off.data(), 8 * (size_t)(f->T + 1));
    p += 8 * (size_t)(f->T + 1);
    auto put = [&](const void *src, bool have) {
      if (have) std::memcpy(p, src, 4 * (size_t)n);
      else std::memset(p, 0, 4 * (size_t)n);
      p += 4 * (size_t)n;
    };
    put(f->feature.data(), true);
    put(f->threshold.data(), true);
    put(f->left.data(), true);
    put(f->right.data(), true);
    put(f->value.data(), true);
    put(f->nsamp.data(), f->nsamp.size() == (size_t)n);
    put(f->depth.data(), f->depth.size() == (size_t)n);
  });
}

int cta_forest_load(cta_ctx *ctx, const void *buf, uint64_t size, cta_forest **out) {
  return guarded([&] {
    if (!ctx || !buf || !out) fail(CTA_EINVAL, "NULL argument");
    ForestHeader h;
    if (size < sizeof h) fail(CTA_EINVAL, "image of %llu bytes is shorter than its header", (unsigned long long)size);
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, kForestMagic, 8) != 0) fail(CTA_EINVAL, "not a serialised forest (bad magic)");
    if (h.version != CTA_FOREST_FORMAT)
      fail(CTA_EUNSUPPORTED, "forest format version %u is not supported (this library reads version %u)",
           h.version, (uint32_t)CTA_FOREST_FORMAT);
    if (h.header_bytes != kForestHeader || h.T <= 0 || h.nodes <= 0 || h.nodes > (int64_t)UINT32_MAX)
      fail(CTA_EINVAL, "corrupt forest header");
    if (forest_image_bytes(h.T, h.nodes) != size)
      fail(CTA_EINVAL, "image holds %llu bytes, its header describes %llu", (unsigned long long)size,
           (unsigned long long)forest_image_bytes(h.T, h.nodes));
    const char *p = static_cast<const char *>(buf) + sizeof h;
    // copies: the arrays of the image need not be aligned
    std::vector<int64_t> off((size_t)h.T + 1);
    std::memcpy(off.data(), p, 8 * off.size());
    p += 8 * off.size();
    if (off[0] != 0 || off[h.T] != h.nodes) fail(CTA_EINVAL, "tree offsets do not match the node count");
    const size_t n = (size_t)h.nodes;
    std::vector<uint32_t> arr(7 * n);
    std::memcpy(arr.data(), p, 28 * n);
    const uint32_t *a = arr.data();
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    *out = import_arrays(ctx, h.T, h.F, h.K, off.data(), reinterpret_cast<const int32_t *>(a),
                         reinterpret_cast<const float *>(a + n), a + 2 * n, a + 3 * n, a + 4 * n, a + 5 * n,
                         a + 6 * n);
  });
}



// Clears the vote matrix v (R x K, device memory), traverses every tree for
// every row and writes the majority label per row; allocates nothing.
static void predict_launch(cta_ctx *ctx, const cta_forest *f, const float *X, int64_t R, uint32_t *labels,
                           uint32_t *v, bool timed) {
  hipStream_t st = ctx->stream;
  const size_t vk = (size_t)R * f->K;
  k_fill_u32<<<grid_for(vk), 256, 0, st>>>(v, vk, 0u);
  {
    std::optional<Timed> tm;
    if (timed) tm.emplace(ctx, "k_predict");
    if (f->T >= (int32_t)kXcd) {
      // trees split over the XCDs: each L2 holds its share of the blocks
      // (faster at every forest size measured, 50 to 500 trees)
      const uint32_t g = grid_for((size_t)R * f->T, 256, 65535u * 4);
      k_predict_xcd<<<(g + kXcd - 1) / kXcd * kXcd, 256, 0, st>>>(
          f->blocks.as<BNode>(), f->broots.as<uint32_t>(), (uint32_t)f->T, X, (uint64_t)R, (uint32_t)f->F,
          (uint32_t)f->K, v);
    } else {
      k_predict<<<grid_for((size_t)R * f->T, 256, 65535u * 4), 256, 0, st>>>(
          f->blocks.as<BNode>(), f->broots.as<uint32_t>(), (uint32_t)f->T, X, (uint64_t)R, (uint32_t)f->F,
          (uint32_t)f->K, v);
    }
  }
  k_vote<<<grid_for(R), 256, 0, st>>>(v, (uint64_t)R, (uint32_t)f->K, labels);
  HIPC(hipGetLastError());
}

static void predict_dev(cta_ctx *ctx, cta_forest *f, const float *X, int64_t R, uint32_t *labels, uint32_t *votes) {
  upload_forest(f);
  uint32_t *v = votes ? votes : ctx->ws_as<uint32_t>("pred_votes", (size_t)R * f->K);
  predict_launch(ctx, f, X, R, labels, v, true);
}


int cta_predict(cta_ctx *ctx, cta_forest *f, const float *X, int64_t R, int32_t F, uint32_t *labels,
                uint32_t *votes) {
  return guarded([&] {
    if (!ctx || !f || !X || !labels) fail(CTA_EINVAL, "NULL argument");
    if (F != f->F) fail(CTA_EINVAL, "X has %d features, the forest %d", F, f->F);
    if (R < 0) fail(CTA_EINVAL, "R must be >= 0");
    if (R == 0) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    float *dX = ctx->ws_as<float>("pred_x", (size_t)R * F);
    uint32_t *dl = ctx->ws_as<uint32_t>("pred_labels", (size_t)R);
    uint32_t *dv = votes ? ctx->ws_as<uint32_t>("pred_vout", (size_t)R * f->K) : nullptr;
    HIPC(hipMemcpyAsync(dX, X, (size_t)R * F * 4, hipMemcpyHostToDevice, ctx->stream));
    predict_dev(ctx, f, dX, R, dl, dv);
    HIPC(hipMemcpyAsync(labels, dl, (size_t)R * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (votes) HIPC(hipMemcpyAsync(votes, dv, (size_t)R * f->K * 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int cta_predict_device(cta_ctx *ctx, cta_forest *f, const float *X_dev, int64_t R, int32_t F, uint32_t *labels_dev,
                       uint32_t *votes_dev) {
  return guarded([&] {
    if (!ctx || !f || !X_dev || !labels_dev) fail(CTA_EINVAL, "NULL argument");
    if (F != f->F) fail(CTA_EINVAL, "X has %d features, the forest %d", F, f->F);
    if (R < 0) fail(CTA_EINVAL, "R must be >= 0");
    if (R == 0) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    predict_dev(ctx, f, X_dev, R, labels_dev, votes_dev);
    ctx->sync();
  });
}

int cta_predict_device_noalloc(cta_ctx *ctx, cta_forest *f, const float *X_dev, int64_t R, int32_t F,
                               uint32_t *labels_dev) {
  return guarded([&] {
    if (!ctx || !f || !X_dev || !labels_dev) fail(CTA_EINVAL, "NULL argument");
    if (F != f->F) fail(CTA_EINVAL, "X has %d features, the forest %d", F, f->F);
    if (R < 0) fail(CTA_EINVAL, "R must be >= 0");
    if (R == 0) return;
    if ((uint64_t)R * (uint64_t)f->K > (uint64_t)CTA_NOALLOC_VOTE_WORDS)
      fail(CTA_EUNSUPPORTED, "R x K = %lld x %d exceeds the %d vote words of the allocation-free predict", (long long)R,
           f->K, (int)CTA_NOALLOC_VOTE_WORDS);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (f->ctx && f->ctx->device != ctx->device)
      fail(CTA_EINVAL, "the forest lives on device %d, the context on device %d", f->ctx->device, ctx->device);
    if (!f->device_ready)
      fail(CTA_EINVAL, "the forest's node table is not on the device (a fit, cta_predict or cta_predict_device "
                       "puts it there)");
    uint32_t *v = scratch_of(ctx->device);
    if (!v) fail(CTA_EINVAL, "device %d has no allocation-free vote buffer", ctx->device);
    HIPC(hipSetDevice(ctx->device));
    predict_launch(ctx, f, X_dev, R, labels_dev, v, false);
    HIPC(hipStreamSynchronize(ctx->stream));
  });
}

int cta_ctx_release_workspace(cta_ctx *ctx) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    ctx->sync();
    ctx->ws_release();
    if (ctx->lane) ctx->lane->ws_release();  // idle between fits
  });
}

int cta_ctx_set_level_hook(cta_ctx *ctx, void (*fn)(void *), void *arg) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "NULL context");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->level_hook = fn;
    ctx->level_hook_arg = arg;
  });
}

int cta_ctx_workspace_bytes(cta_ctx *ctx, uint64_t *bytes) {
  return guarded([&] {
    if (!ctx || !bytes) fail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *bytes = ctx->ws_bytes();
  });
}

int cta_ctx_set_workspace_limit(cta_ctx *ctx, uint64_t bytes) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->ws_limit = bytes;
  });
}

int cta_fit_estimate(int32_t N, int32_t F, int32_t K, int32_t T, const cta_fit_opts *opts, uint64_t *dataset_bytes,
                     uint64_t *workspace_bytes) {
  return guarded([&] {
    if (!opts || !dataset_bytes || !workspace_bytes) fail(CTA_EINVAL, "NULL argument");
    check_sizes(N, F, K);
    if (T <= 0) fail(CTA_EINVAL, "T must be positive");
    const uint64_t n = (uint64_t)N, f = (uint64_t)F, nf = n * f;
    const size_t b = N <= 65535 ? 2 : 4;
    const uint64_t lab = K <= 255 ? 1 : 2;
    // the dataset's own buffers: X_T, labels, presorted rows, code columns
    // (at most 16 bits per value), row-free signatures and feature tables,
    // plus the copies of host X and y while it is built
    *dataset_bytes = nf * 4 + n * lab + nf * b + 2 * nf + 256 * f + 4 * n * ((f + 31) / 32) + 64 * f + 4096 +
                     nf * 4 + n * 4;
    // the context's workspaces: radix-sort and code-column temporaries of the
    // build, then per resident tree (every feature keeping rows), the node
    // arena and canonical export of at most two nodes per in-bag sample, and
    // the device predict table
    const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
    const uint64_t build = 6 * nf * 4 + f * kDigits * ntiles * 4 + 64 * f + ((n + kCodeTile - 1) / kCodeTile) * f * 4;
    const int32_t m = opts->max_features > 0 ? opts->max_features : (int32_t)std::ceil(std::sqrt((double)F));
    const int64_t Tb = opts->mode != CTA_MODE_BATCHED ? 1
                       : opts->trees_per_batch > 0 ? std::min(opts->trees_per_batch, T)
                                                   : T;
    const uint64_t per_tree = tree_bytes(N, F, K, m, b) + 2 * n * (13 * 4 + sizeof(PNode));
    *workspace_bytes = build + (uint64_t)Tb * per_tree + (uint64_t)T * 2 * n * sizeof(PNode) + (64ull << 20);
  });
}

int cta_ctx_set_profiling(cta_ctx *ctx, int enable) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->sync();
    ctx->profiling = enable != 0;
  });
}

int cta_ctx_reset_stats(cta_ctx *ctx) {
  return guarded([&] {
    if (!ctx) fail(CTA_EINVAL, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->sync();
    ctx->stats.clear();
  });
}

int cta_ctx_kernel_stats(cta_ctx *ctx, cta_kernel_stat *out, int32_t max, int32_t *n) {
  return guarded([&] {
    if (!ctx || !n) fail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->sync();
    int32_t i = 0;
    for (auto &kv : ctx->stats) {
      if (out && i < max) {
        cta_kernel_stat &s = out[i];
        std::memset(&s, 0, sizeof(s));
        std::strncpy(s.name, kv.first.c_str(), sizeof(s.name) - 1);
        s.launches = kv.second.launches;
        s.total_ms = kv.second.ms;
        s.alg_bytes = kv.second.bytes;
      }
      ++i;
    }
    *n = i;
  });
}

}  // extern "C"

// ---- per-kernel parity hooks --------------------------------------------------------
// Row entry of sample s as the kernels store it (label in the high bits of
// 32-bit entries).
template <class IdxT, class LabT>
static IdxT row_entry(uint32_t s, const std::vector<LabT> &y) {
  if constexpr (sizeof(IdxT) == 4) return (IdxT)(s | ((uint32_t)y[s] << kIdBits));
  else return (IdxT)s;
}

template <class IdxT, class LabT>
static void hook_split_scan(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const int32_t *features,
                            int32_t m_eval, int32_t chunk, int32_t *bf, uint32_t *bp, uint64_t *num, uint64_t *den) {
  const int32_t N = ds->N, F = ds->F, K = ds->K;
  hipStream_t st = ctx->stream;
  const uint32_t Cs = chunk > 0 ? (uint32_t)chunk : kScanChunk;
  // node rows into tree 0's buffer 0
  std::vector<IdxT> hrows((size_t)F * N, 0);
  std::vector<uint32_t> counts(K, 0);
  std::vector<LabT> hy(N);
  HIPC(hipMemcpyAsync(hy.data(), ds->y.p, (size_t)N * sizeof(LabT), hipMemcpyDeviceToHost, st));
  ctx->sync();
  for (int32_t f = 0; f < F; ++f)
    for (uint32_t i = 0; i < n; ++i) hrows[(size_t)f * N + i] = row_entry<IdxT>(rows[(size_t)f * n + i], hy);
  for (uint32_t i = 0; i < n; ++i) counts[hy[rows[i]]]++;
  const uint32_t nch = (n + Cs - 1) / Cs;
  const uint32_t fpt = nch > 1 ? 1u : std::max(1u, Cs / n);
  const uint32_t nf = (uint32_t)m_eval;
  const uint32_t ntask = nf == 0 ? 0u : (nch > 1 ? nf * nch : (nf + fpt - 1) / fpt);
  std::vector<Task> tasks(std::max(ntask, 1u));
  for (uint32_t t = 0; t < ntask; ++t) {
    if (nch > 1) tasks[t] = Task{0u, (t / nch) | (1u << 16), t % nch, nch};
    else {
      const uint32_t first = t * fpt;
      tasks[t] = Task{0u, first | (std::min(fpt, nf - first) << 16), 0u, nch};
    }
  }
  NAux ax = {};
  ax.scan_base = 0;
  ax.scan_ntask = ntask;
  ax.nch_s = nch;
  ax.fpt = fpt;
  ax.hist_base = 0;
  ax.best_f = -1;
  ax.rl_feat = -1;  // no row-free features here
  ax.nf = nf;
  const uint32_t Cf = kFillChunk;
  ax.fill_base = 0;
  ax.nch_f = (n + Cf - 1) / Cf;
  FNode nd = {0u, 0u, n, 0u, 0u, 1u, 0u, 0u};
  Ctl hc = {};
  hc.q = 1;
  hc.n_scan = ntask;
  hc.n_cls[0] = ntask;  // one task list, class 0
  hc.n_nodes = 1;
  const size_t nbslot = kByteKinds * (size_t)kByteSlots * kByteStride;
  DevMem d_rows(hrows.size() * sizeof(IdxT)), d_front(sizeof(FNode)), d_cnt(K * 4), d_aux(sizeof(NAux)),
      d_feats(std::max(nf, 1u) * 4), d_tasks(tasks.size() * sizeof(Task)),
      d_res(std::max(ntask, 1u) * sizeof(ScanRes)), d_hist((size_t)std::max(nf, 1u) * nch * K * 4), d_ctl(sizeof(Ctl)),
      d_cntL(K * 4), d_ar(7 * 4 * 4), d_fill(((size_t)ax.nch_f + 1) * sizeof(Task)), d_bslot(nbslot * 8),
      d_vals(kChFields * 4);
  k_fill_u32<<<grid_for(nbslot * 2), 256, 0, st>>>(d_bslot.as<uint32_t>(), nbslot * 2, 0u);
  HIPC(hipMemcpyAsync(d_rows.p, hrows.data(), hrows.size() * sizeof(IdxT), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_front.p, &nd, sizeof nd, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_cnt.p, counts.data(), K * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_aux.p, &ax, sizeof ax, hipMemcpyHostToDevice, st));
  if (nf) HIPC(hipMemcpyAsync(d_feats.p, features, nf * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(Task), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_ctl.p, &hc, sizeof hc, hipMemcpyHostToDevice, st));
  FitArgs<IdxT, LabT> a = {};
  a.X_T = ds->XT.as<float>();
  a.y = ds->y.as<LabT>();
  a.si0 = ds->si0.as<IdxT>();
  a.codes = ds->codes.as<uint8_t>();
  a.fcode = ds->fcode.as<uint64_t>();
  a.Fr = (uint32_t)ds->F;  // every feature keeps its row
  a.slot_of = ds->ident.as<int32_t>();
  a.row_feat = ds->ident.as<int32_t>();
  a.rl_of = ds->rl_none.as<int32_t>();
  a.buf0 = d_rows.as<IdxT>();
  a.buf1 = d_rows.as<IdxT>();
  a.N = N;
  a.F = F;
  a.K = K;
  a.m = (int32_t)std::max(nf, 1u);
  a.Cs = Cs;
  a.Cf = Cf;
  a.Cr = 8192;
  a.front = d_front.as<FNode>();
  a.cnt = d_cnt.as<uint32_t>();
  a.aux = d_aux.as<NAux>();
  a.feats = d_feats.as<int32_t>();
  a.scan_tasks = d_tasks.as<Task>();
  a.res = d_res.as<ScanRes>();
  a.hist = d_hist.as<uint32_t>();
  a.cntL = d_cntL.as<uint32_t>();
  a.fill_tasks = d_fill.as<Task>();
  a.vals = d_vals.as<uint32_t>();  // the children-scan counts k_decide writes
  a.bslot = d_bslot.as<unsigned long long>();
  uint32_t *ar = d_ar.as<uint32_t>();
  a.arena = Arena{(int32_t *)ar, (float *)(ar + 4), ar + 8, ar + 12, ar + 16, ar + 20, ar + 24};
  a.ctl = d_ctl.as<Ctl>();
  a.cap_q = 1;
  a.cap_scan = std::max(ntask, 1u);
  a.cap_hist = std::max(nf, 1u) * nch * K;
  a.cap_fill = ax.nch_f + 1;
  a.cap_nodes = 3;
  const size_t sm = (size_t)kWavesPerBlock * 2 * K * 4;
  const uint32_t G = 64;
  k_scan_hist<IdxT, LabT><<<G, kBlock, sm, st>>>(a);
  k_scan_carry<IdxT, LabT><<<G, kBlock, 0, st>>>(a);
  if ((uint32_t)K <= kScanRegsK && !a.hist_lds) k_scan<IdxT, LabT, true><<<G, kBlock, sm, st>>>(a);
  else k_scan<IdxT, LabT, false><<<G, kBlock, sm, st>>>(a);
  k_decide<IdxT, LabT><<<G, kBlock, sm, st>>>(a);
  HIPC(hipGetLastError());
  NAux out = {};
  std::vector<ScanRes> res(std::max(ntask, 1u));
  HIPC(hipMemcpyAsync(&out, d_aux.p, sizeof out, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(res.data(), d_res.p, res.size() * sizeof(ScanRes), hipMemcpyDeviceToHost, st));
  ctx->sync();
  *bf = out.best_f;
  *bp = out.best_p;
  *num = 0;
  *den = 0;
  for (uint32_t t = 0; t < ntask; ++t)
    if (res[t].ok && res[t].feat == out.best_f && res[t].pos == out.best_p) {
      *num = res[t].num;
      *den = res[t].den;
    }
}

extern "C" int cta_k_split_scan(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const int32_t *features,
                     int32_t m_eval, int32_t chunk, int32_t *best_feature, uint32_t *best_pos, uint64_t *num,
                     uint64_t *den) {
  return guarded([&] {
    if (!ctx || !ds || !rows || !best_feature || !best_pos || !num || !den) fail(CTA_EINVAL, "NULL argument");
    if (n < 1 || n > (uint32_t)ds->N) fail(CTA_EINVAL, "n out of range");
    if (m_eval < 0 || m_eval > ds->F || (m_eval > 0 && !features)) fail(CTA_EINVAL, "bad feature list");
    if (chunk != 0 && chunk < 64) fail(CTA_EINVAL, "chunk must be 0 or >= 64");
    for (int32_t j = 0; j < m_eval; ++j)
      if (features[j] < 0 || features[j] >= ds->F) fail(CTA_EINVAL, "feature out of range");
    for (size_t i = 0; i < (size_t)ds->F * n; ++i)
      if (rows[i] >= (uint32_t)ds->N) fail(CTA_EINVAL, "row index out of range");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    const bool b16 = ds->idx_bits == 16, l8 = ds->lab_bits == 8;
    if (b16 && l8) hook_split_scan<uint16_t, uint8_t>(ctx, ds, rows, n, features, m_eval, chunk, best_feature, best_pos, num, den);
    else if (b16) hook_split_scan<uint16_t, uint16_t>(ctx, ds, rows, n, features, m_eval, chunk, best_feature, best_pos, num, den);
    else if (l8) hook_split_scan<uint32_t, uint8_t>(ctx, ds, rows, n, features, m_eval, chunk, best_feature, best_pos, num, den);
    else hook_split_scan<uint32_t, uint16_t>(ctx, ds, rows, n, features, m_eval, chunk, best_feature, best_pos, num, den);
  });
}

template <class IdxT, class LabT>
static void hook_reshuffle(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const uint8_t *mark,
                           int32_t chunk, int32_t global_bits, uint32_t *out) {
  const int32_t N = ds->N, F = ds->F, K = ds->K;
  hipStream_t st = ctx->stream;
  const uint32_t Cr = chunk > 0 ? (uint32_t)chunk : kReshChunk;
  const uint32_t W = ((uint32_t)N + 31u) / 32u;
  std::vector<LabT> hy(N);
  HIPC(hipMemcpyAsync(hy.data(), ds->y.p, (size_t)N * sizeof(LabT), hipMemcpyDeviceToHost, st));
  ctx->sync();
  std::vector<IdxT> hrows((size_t)F * N, 0);
  for (int32_t f = 0; f < F; ++f)
    for (uint32_t i = 0; i < n; ++i) hrows[(size_t)f * N + i] = row_entry<IdxT>(rows[(size_t)f * n + i], hy);
  std::vector<uint32_t> hbits(W, 0);
  for (int32_t s = 0; s < N; ++s)
    if (mark[s]) hbits[s >> 5] |= 1u << (s & 31);
  uint32_t nL = 0;
  for (uint32_t i = 0; i < n; ++i) nL += mark[rows[i]] ? 1 : 0;
  // the children-scan counts of k_decide for one node of tree 0, whose frontier
  // range is [0, 1) and whose tasks start at offset 0
  NAux ax = {};
  ax.nL = nL;
  ax.flags = kSplit | kLiveL | kLiveR;
  ax.dst = 2;
  uint32_t v[kChFields] = {0u, 0u, 0u, 0u, 0u};
  {
    const bool big = n >= kBigN && (uint32_t)F < kBigMaxRows;
    const uint32_t rpt = big ? 1u : std::max(1u, Cr / n);
    const uint32_t ntask = big ? (uint32_t)F : ((uint32_t)F + rpt - 1) / rpt;
    ax.resh_big = big ? 1u : 0u;
    ax.rpt = rpt;
    ax.resh_ntask = ntask;
    v[2] = big ? (uint32_t)F : 0u;
    v[3] = big ? 0u : ntask;
    v[4] = big ? (uint32_t)F * ((n + kReshTile - 1) / kReshTile) : 0u;
  }
  FNode nd = {0u, 0u, n, 0u, 0u, 1u, 0u, 0u};
  Ctl hc = {};
  hc.q = 1;
  hc.c_big = v[2];
  hc.c_small = v[3];
  hc.c_tiles = v[4];
  const uint32_t zeros[kChFields] = {0u, 0u, 0u, 0u, 0u};
  const uint32_t range[2] = {0u, 1u};
  DevMem d_in(hrows.size() * sizeof(IdxT)), d_out(hrows.size() * sizeof(IdxT)), d_bits((size_t)W * 4),
      d_front(sizeof(FNode)), d_aux(sizeof(NAux)), d_big((size_t)F * sizeof(Task)),
      d_small((size_t)std::max(v[3], 1u) * sizeof(Task)), d_offs(sizeof zeros), d_range(sizeof range), d_plan(4 * 4),
      d_wg(4 * 4), d_ctl(sizeof(Ctl));
  HIPC(hipMemcpyAsync(d_in.p, hrows.data(), hrows.size() * sizeof(IdxT), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_out.p, hrows.data(), hrows.size() * sizeof(IdxT), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_bits.p, hbits.data(), (size_t)W * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_front.p, &nd, sizeof nd, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_aux.p, &ax, sizeof ax, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_offs.p, zeros, sizeof zeros, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_range.p, range, sizeof range, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(d_ctl.p, &hc, sizeof hc, hipMemcpyHostToDevice, st));
  FitArgs<IdxT, LabT> a = {};
  a.X_T = ds->XT.as<float>();
  a.y = ds->y.as<LabT>();
  a.si0 = ds->si0.as<IdxT>();
  a.codes = ds->codes.as<uint8_t>();
  a.fcode = ds->fcode.as<uint64_t>();
  a.Fr = (uint32_t)ds->F;  // every feature keeps its row
  a.slot_of = ds->ident.as<int32_t>();
  a.row_feat = ds->ident.as<int32_t>();
  a.rl_of = ds->rl_none.as<int32_t>();
  a.buf0 = d_in.as<IdxT>();
  a.buf1 = d_out.as<IdxT>();
  a.bits = d_bits.as<uint32_t>();
  a.W = W;
  a.T = 1;
  a.N = N;
  a.F = F;
  a.K = K;
  a.Cr = Cr;
  a.front = d_front.as<FNode>();
  a.aux = d_aux.as<NAux>();
  a.big_tasks = d_big.as<Task>();
  a.small_tasks = d_small.as<Task>();
  a.offs = d_offs.as<uint32_t>();
  a.tlo = d_range.as<uint32_t>();
  a.thi = d_range.as<uint32_t>() + 1;
  a.plan = d_plan.as<uint32_t>();
  a.wg_off = d_wg.as<uint32_t>();
  a.ctl = d_ctl.as<Ctl>();
  a.cap_big = (uint32_t)F;
  a.cap_small = std::max(v[3], 1u);
  k_emit_resh<IdxT, LabT><<<8, kBlock, 0, st>>>(a);
  k_plan<IdxT, LabT><<<1, kPlanThreads, 0, st>>>(a, 16u, 0u);
  launch_resh(ctx, a, global_bits != 0, 16u);
  HIPC(hipGetLastError());
  std::vector<IdxT> hout(hrows.size());
  Ctl back = {};
  HIPC(hipMemcpyAsync(hout.data(), d_out.p, hout.size() * sizeof(IdxT), hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(&back, d_ctl.p, sizeof back, hipMemcpyDeviceToHost, st));
  ctx->sync();
  if (back.err) fail(CTA_EINTERNAL, "reshuffle hook: device capacity check failed");
  for (int32_t f = 0; f < F; ++f)
    for (uint32_t i = 0; i < n; ++i) out[(size_t)f * n + i] = sample_of(hout[(size_t)f * N + i]);
}

extern "C" int cta_k_reshuffle(cta_ctx *ctx, cta_dataset *ds, const uint32_t *rows, uint32_t n, const uint8_t *mark,
                               int32_t chunk, int32_t global_bits, uint32_t *out) {
  return guarded([&] {
    if (!ctx || !ds || !rows || !mark || !out) fail(CTA_EINVAL, "NULL argument");
    if (n < 1 || n > (uint32_t)ds->N) fail(CTA_EINVAL, "n out of range");
    if (chunk != 0 && chunk < 64) fail(CTA_EINVAL, "chunk must be 0 or >= 64");
    for (size_t i = 0; i < (size_t)ds->F * n; ++i)
      if (rows[i] >= (uint32_t)ds->N) fail(CTA_EINVAL, "row index out of range");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPC(hipSetDevice(ctx->device));
    const bool b16 = ds->idx_bits == 16, l8 = ds->lab_bits == 8;
    if (b16 && l8) hook_reshuffle<uint16_t, uint8_t>(ctx, ds, rows, n, mark, chunk, global_bits, out);
    else if (b16) hook_reshuffle<uint16_t, uint16_t>(ctx, ds, rows, n, mark, chunk, global_bits, out);
    else if (l8) hook_reshuffle<uint32_t, uint8_t>(ctx, ds, rows, n, mark, chunk, global_bits, out);
    else hook_reshuffle<uint32_t, uint16_t>(ctx, ds, rows, n, mark, chunk, global_bits, out);
  });
}

// serve.hip -- serving loop of libcudatree_amd.so: resident forests, predict
// requests coalesced into shared traversal launches, fit jobs admitted by HBM
// accounting, per-request latency.  Built only on the public C-ABI
// (include/cudatree_amd.h) plus HIP runtime calls for its staging buffers.
//
// The reference has no serving path: every predict call re-uploads the rows
// and every tree (base_tree.py:57-62) and votes in a Python loop
// (random_forest.py:355-358).  Here forests stay resident, a predict worker
// (own context, own stream) drains the request queue batch by batch --
// requests that arrive while a batch is on the device join the next one --
// and fit workers (own contexts and streams) grow new forests concurrently,
// each admitted only when its estimated device footprint fits the budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cudatree_amd.h"

namespace {

thread_local std::string s_err;

struct SError : std::runtime_error {
  int code;
  SError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void sfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw SError(code, buf);
}

void cta_ok(int rc, const char *what) {
  if (rc != CTA_OK) sfail(rc, "%s: %s", what, cta_last_error());
}

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) sfail(CTA_EHIP, "%s: %s", what, hipGetErrorString(e));
}

template <class F>
int sguard(F &&f) {
  try {
    f();
    s_err.clear();
    return CTA_OK;
  } catch (const SError &e) {
    s_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    s_err = "host allocation failed";
    return CTA_ENOMEM;
  } catch (const std::exception &e) {
    s_err = e.what();
    return CTA_EINTERNAL;
  }
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device bytes a fit job is admitted with: the library's estimate of its
// dataset and of the workspaces it may hold (the fit context is capped at
// the latter, so the job cannot grow past it).
void fit_estimate(int32_t N, int32_t F, int32_t K, int32_t T, const cta_fit_opts &o, uint64_t *ds, uint64_t *ws) {
  cta_ok(cta_fit_estimate(N, F, K, T, &o, ds, ws), "cta_fit_estimate");
}

}  // namespace

struct PredReq {
  uint64_t ticket;
  int32_t forest;
  const float *X;
  int64_t R;
  uint32_t *out;
};

struct FitJob {
  uint64_t ticket;
  std::vector<float> X;
  std::vector<uint32_t> y;
  int32_t N, F, K, T;
  cta_fit_opts opts;
  uint64_t est;     // admitted device bytes (dataset + workspace cap)
  uint64_t ws_cap;  // workspace cap of the fit context
};

struct TicketState {
  int32_t status = -1;  // -1 pending, else a CTA_* status
  int32_t forest = -1;
  double t_submit = 0, t_done = 0;
};

struct ForestEntry {
  cta_forest *forest = nullptr;
  int32_t T = 0, F = 0, K = 0;
  uint64_t bytes = 0;
  int refs = 0;
  bool dropped = false;
};

// One predict lane: its context (stream), staging buffers and worker thread.
// With coalescing, lane 0 serves the requests of at most split_rows rows and
// lane 1 the larger ones, so short requests never queue behind the traversal
// of a long one (the two lanes' launches run side by side on the device).
struct Lane {
  int cls = 0;  // 0: short requests (or every request with one lane), 1: long
  cta_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  float *hX = nullptr, *dX = nullptr;
  uint32_t *hL = nullptr, *dL = nullptr;
  size_t cap_rows = 0, cap_F = 0;
  bool busy = false;
  std::thread th;
};

struct cta_server {
  cta_server_opts o{};
  uint64_t budget = 0, in_use = 0, peak = 0;
  std::vector<std::unique_ptr<Lane>> lanes;
  int64_t split_rows = 0;  // largest short request (0: one lane for every request)
  std::vector<cta_ctx *> fctx;
  std::mutex mu;
  std::condition_variable cv_pred, cv_fit, cv_done;
  std::condition_variable cv_idle;  // predict queue empty and no batch on the device
  int busy_lanes = 0;
  std::deque<PredReq> pq[2];  // queued predict requests of each lane class, in arrival order
  std::deque<FitJob> fq;
  std::map<int32_t, ForestEntry> forests;
  std::unordered_map<uint64_t, TicketState> tickets;
  uint64_t next_ticket = 1;
  int32_t next_forest = 0;
  bool stop = false;
  std::vector<std::thread> fthreads;
  // statistics
  std::vector<double> lat_ms;
  uint64_t pred_requests = 0, pred_rows = 0, pred_launches = 0, fit_done = 0, fit_rejected = 0;
  double t_first = 0, t_last = 0;
  // device free-memory drop over each fit job (dataset build to the end of
  // its fit), against the bytes it was admitted with
  uint64_t fit_mem_max = 0;
  double fit_mem_ratio = 0;

  cta_ctx *home_ctx() const { return lanes[0]->ctx; }

  void register_forest_locked(cta_forest *f, int32_t *id) {
    int32_t T = 0, F = 0, K = 0;
    int64_t n = 0;
    cta_ok(cta_forest_info(f, &T, &n, &F, &K), "cta_forest_info");
    ForestEntry e;
    e.forest = f;
    e.T = T;
    e.F = F;
    e.K = K;
    e.bytes = (uint64_t)n * 16 + (uint64_t)(T + 1) * 4;
    in_use += e.bytes;
    peak = std::max(peak, in_use);
    *id = next_forest++;
    forests[*id] = e;
  }

  void release_forest_locked(int32_t id) {
    auto it = forests.find(id);
    if (it == forests.end() || it->second.refs > 0 || !it->second.dropped) return;
    in_use -= it->second.bytes;
    cta_forest_destroy(it->second.forest);
    forests.erase(it);
  }

  void ensure_staging(Lane &L, int64_t rows, int32_t F) {
    const size_t need = (size_t)rows * F;
    if (L.cap_rows * L.cap_F >= need && L.cap_rows >= (size_t)rows) return;
    const size_t r =
        std::max<size_t>((size_t)rows, L.cap_rows * 2 > (size_t)rows ? L.cap_rows * 2 : (size_t)rows);
    const size_t f = std::max<size_t>((size_t)F, L.cap_F);
    if (L.hX) (void)hipHostFree(L.hX);
    if (L.hL) (void)hipHostFree(L.hL);
    if (L.dX) (void)hipFree(L.dX);
    if (L.dL) (void)hipFree(L.dL);
    L.hX = nullptr;
    L.hL = nullptr;
    L.dX = nullptr;
    L.dL = nullptr;
    {
      std::lock_guard<std::mutex> lk(mu);
      in_use -= (uint64_t)L.cap_rows * L.cap_F * 4 + (uint64_t)L.cap_rows * 4;
      L.cap_rows = 0;
      L.cap_F = 0;
      in_use += (uint64_t)r * f * 4 + (uint64_t)r * 4;
      peak = std::max(peak, in_use);
    }
    hip_ok(hipHostMalloc((void **)&L.hX, r * f * 4, hipHostMallocDefault), "hipHostMalloc");
    hip_ok(hipHostMalloc((void **)&L.hL, r * 4, hipHostMallocDefault), "hipHostMalloc");
    hip_ok(hipMalloc((void **)&L.dX, r * f * 4), "hipMalloc");
    hip_ok(hipMalloc((void **)&L.dL, r * 4), "hipMalloc");
    L.cap_rows = r;
    L.cap_F = f;
  }

  // Lane of a request of R rows.
  int lane_of(int64_t R) const { return (lanes.size() > 1 && R > split_rows) ? 1 : 0; }

  bool queues_empty() const { return pq[0].empty() && pq[1].empty(); }

  // One coalesced (or single) launch over the requests at the front of the
  // lane's queue: the run of requests of the first one's forest, in arrival
  // order.
  void predict_loop(Lane &L) {
    (void)hipSetDevice(o.device);
    const int64_t max_rows = o.max_batch_rows > 0 ? o.max_batch_rows : (1 << 20);
    const int64_t max_work = o.max_batch_work > 0 ? o.max_batch_work : (int64_t)1 << 22;
    for (;;) {
      std::vector<PredReq> batch;
      ForestEntry fe;
      int32_t fid = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        std::deque<PredReq> &q = pq[L.cls];
        cv_pred.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;  // stopping, nothing left for this lane
        fid = q.front().forest;
        auto fit = forests.find(fid);
        const int64_t T = fit != forests.end() ? std::max<int64_t>(1, fit->second.T) : 1;
        int64_t rows = 0;
        // a batch stays under max_rows rows and max_work row-tree visits (its
        // first request is always taken)
        while (!q.empty() && q.front().forest == fid &&
               (batch.empty() || (rows + q.front().R <= max_rows && (rows + q.front().R) * T <= max_work))) {
          rows += q.front().R;
          batch.push_back(q.front());
          q.pop_front();
          if (!o.coalesce) break;
        }
        auto f = forests.find(fid);
        if (f == forests.end()) {
          for (auto &r : batch) finish_locked(r.ticket, CTA_EINVAL, -1);
          continue;
        }
        f->second.refs++;
        fe = f->second;
        L.busy = true;
        ++busy_lanes;
      }
      int rc = CTA_OK;
      int64_t rows = 0;
      for (auto &r : batch) rows += r.R;
      try {
        ensure_staging(L, rows, fe.F);
        int64_t off = 0;
        for (auto &r : batch) {
          std::memcpy(L.hX + (size_t)off * fe.F, r.X, (size_t)r.R * fe.F * 4);
          off += r.R;
        }
        hip_ok(hipMemcpyAsync(L.dX, L.hX, (size_t)rows * fe.F * 4, hipMemcpyHostToDevice, L.stream), "copy in");
        hip_ok(hipStreamSynchronize(L.stream), "sync");
        rc = cta_predict_device(L.ctx, fe.forest, L.dX, rows, fe.F, L.dL, nullptr);
        if (rc == CTA_OK) {
          hip_ok(hipMemcpyAsync(L.hL, L.dL, (size_t)rows * 4, hipMemcpyDeviceToHost, L.stream), "copy out");
          hip_ok(hipStreamSynchronize(L.stream), "sync");
          off = 0;
          for (auto &r : batch) {
            std::memcpy(r.out, L.hL + off, (size_t)r.R * 4);
            off += r.R;
          }
        }
      } catch (const SError &e) {
        rc = e.code;
      }
      std::lock_guard<std::mutex> lk(mu);
      pred_launches += 1;
      for (auto &r : batch) {
        pred_requests += 1;
        pred_rows += (uint64_t)r.R;
        finish_locked(r.ticket, rc, -1);
      }
      forests[fid].refs--;
      release_forest_locked(fid);
      L.busy = false;
      --busy_lanes;
      if (queues_empty() && busy_lanes == 0) cv_idle.notify_all();
    }
  }

  // Runs on a fit worker between groups of levels: queued or running
  // predict batches get the device first, for at most fit_yield_us.
  static void yield_to_predict(void *arg) {
    cta_server *s = static_cast<cta_server *>(arg);
    const int32_t us = s->o.fit_yield_us;
    if (us < 0) return;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_idle.wait_for(lk, std::chrono::microseconds(us ? us : 4000),
                        [&] { return s->stop || (s->queues_empty() && s->busy_lanes == 0); });
  }

  void fit_loop(size_t w) {
    (void)hipSetDevice(o.device);
    (void)cta_ctx_set_level_hook(fctx[w], &cta_server::yield_to_predict, this);
    for (;;) {
      FitJob job;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (stop && fq.empty()) return;
          if (!fq.empty()) {
            FitJob &h = fq.front();
            if (h.est > budget) {  // can never fit: reject
              const uint64_t t = h.ticket;
              fq.pop_front();
              fit_rejected += 1;
              finish_locked(t, CTA_ENOMEM, -1);
              continue;
            }
            if (in_use + h.est <= budget) break;  // admit
          }
          cv_fit.wait(lk);
        }
        job = std::move(fq.front());
        fq.pop_front();
        in_use += job.est;
        peak = std::max(peak, in_use);
      }
      int rc = CTA_OK;
      cta_forest *forest = nullptr;
      cta_dataset *ds = nullptr;
      uint64_t free0 = 0, free1 = 0, tot = 0;
      (void)cta_ctx_mem_info(fctx[w], &free0, &tot);
      rc = cta_ctx_set_workspace_limit(fctx[w], job.ws_cap);
      if (rc == CTA_OK) rc = cta_dataset_create(fctx[w], job.X.data(), job.y.data(), job.N, job.F, job.K, 0, &ds);
      if (rc == CTA_OK) rc = cta_fit(fctx[w], ds, job.T, nullptr, nullptr, &job.opts, &forest);
      (void)cta_ctx_mem_info(fctx[w], &free1, &tot);
      if (ds) cta_dataset_destroy(ds);
      cta_ctx_release_workspace(fctx[w]);
      std::lock_guard<std::mutex> lk(mu);
      if (rc == CTA_OK && free0 > free1) {
        fit_mem_max = std::max(fit_mem_max, free0 - free1);
        fit_mem_ratio = std::max(fit_mem_ratio, (double)(free0 - free1) / (double)job.est);
      }
      in_use -= job.est;
      int32_t id = -1;
      if (rc == CTA_OK) {
        try {
          register_forest_locked(forest, &id);
        } catch (const SError &e) {
          rc = e.code;
          cta_forest_destroy(forest);
        }
      }
      fit_done += rc == CTA_OK ? 1 : 0;
      finish_locked(job.ticket, rc, id);
      cv_fit.notify_all();
    }
  }

  void finish_locked(uint64_t ticket, int32_t status, int32_t forest) {
    auto it = tickets.find(ticket);
    if (it == tickets.end()) return;
    it->second.status = status;
    it->second.forest = forest;
    it->second.t_done = now_ms();
    t_last = it->second.t_done;
    if (forest < 0 && status == CTA_OK) lat_ms.push_back(it->second.t_done - it->second.t_submit);
    cv_done.notify_all();
  }

  uint64_t new_ticket_locked() {
    const uint64_t t = next_ticket++;
    TicketState s;
    s.t_submit = now_ms();
    if (t_first == 0) t_first = s.t_submit;
    tickets[t] = s;
    return t;
  }
};

extern "C" {

const char *cta_server_last_error(void) { return s_err.c_str(); }

void cta_server_opts_default(cta_server_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->device = 0;
  o->coalesce = 1;
  o->max_batch_rows = 1 << 20;
  o->max_batch_work = (int64_t)1 << 22;
  o->hbm_budget = 0;
  o->fit_workers = 1;
}

int cta_server_create(const cta_server_opts *opts, cta_server **out) {
  return sguard([&] {
    if (!opts || !out) sfail(CTA_EINVAL, "NULL argument");
    std::unique_ptr<cta_server> s(new cta_server());
    s->o = *opts;
    if (s->o.fit_workers <= 0) s->o.fit_workers = 1;
    // a second (long-request) lane only when coalescing; split_rows < 0: one lane
    s->split_rows = (s->o.coalesce && s->o.split_rows >= 0) ? (s->o.split_rows ? s->o.split_rows : 4096) : 0;
    const int nl = s->split_rows > 0 ? 2 : 1;
    for (int l = 0; l < nl; ++l) {
      std::unique_ptr<Lane> L(new Lane());
      L->cls = l;
      cta_ok(cta_ctx_create(s->o.device, &L->ctx), "cta_ctx_create");
      hip_ok(hipSetDevice(s->o.device), "hipSetDevice");
      hip_ok(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
      s->lanes.push_back(std::move(L));
    }
    for (int i = 0; i < s->o.fit_workers; ++i) {
      cta_ctx *c = nullptr;
      cta_ok(cta_ctx_create(s->o.device, &c), "cta_ctx_create");
      s->fctx.push_back(c);
    }
    uint64_t fr = 0, tot = 0;
    cta_ok(cta_ctx_mem_info(s->home_ctx(), &fr, &tot), "cta_ctx_mem_info");
    s->budget = s->o.hbm_budget ? s->o.hbm_budget : (uint64_t)(0.9 * (double)fr);
    cta_server *raw = s.get();
    for (auto &L : raw->lanes) {
      Lane *lp = L.get();
      lp->th = std::thread([raw, lp] { raw->predict_loop(*lp); });
    }
    for (size_t w = 0; w < raw->fctx.size(); ++w) raw->fthreads.emplace_back([raw, w] { raw->fit_loop(w); });
    *out = s.release();
  });
}

int cta_server_destroy(cta_server *s) {
  return sguard([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_pred.notify_all();
    s->cv_fit.notify_all();
    s->cv_idle.notify_all();
    for (auto &L : s->lanes)
      if (L->th.joinable()) L->th.join();
    for (auto &t : s->fthreads)
      if (t.joinable()) t.join();
    for (auto &kv : s->forests) cta_forest_destroy(kv.second.forest);
    (void)hipSetDevice(s->o.device);
    for (auto &L : s->lanes) {
      if (L->hX) (void)hipHostFree(L->hX);
      if (L->hL) (void)hipHostFree(L->hL);
      if (L->dX) (void)hipFree(L->dX);
      if (L->dL) (void)hipFree(L->dL);
      if (L->stream) (void)hipStreamDestroy(L->stream);
      cta_ctx_destroy(L->ctx);
    }
    for (auto c : s->fctx) cta_ctx_destroy(c);
    delete s;
  });
}

int cta_server_add_forest(cta_server *s, const cta_forest *f, int32_t *forest_id) {
  return sguard([&] {
    if (!s || !f || !forest_id) sfail(CTA_EINVAL, "NULL argument");
    int32_t T = 0, F = 0, K = 0;
    int64_t n = 0;
    cta_ok(cta_forest_info(f, &T, &n, &F, &K), "cta_forest_info");
    std::vector<int64_t> off(T + 1);
    std::vector<int32_t> feat(n);
    std::vector<float> thr(n);
    std::vector<uint32_t> l(n), r(n), v(n);
    cta_ok(cta_forest_export(f, off.data(), feat.data(), thr.data(), l.data(), r.data(), v.data(), nullptr, nullptr),
           "cta_forest_export");
    cta_forest *copy = nullptr;
    cta_ok(cta_forest_import(s->home_ctx(), T, F, K, off.data(), feat.data(), thr.data(), l.data(), r.data(), v.data(),
                             &copy),
           "cta_forest_import");
    std::lock_guard<std::mutex> lk(s->mu);
    s->register_forest_locked(copy, forest_id);
  });
}

int cta_server_get_forest(cta_server *s, int32_t forest_id, cta_forest **forest) {
  return sguard([&] {
    if (!s || !forest) sfail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->forests.find(forest_id);
    if (it == s->forests.end() || it->second.dropped) sfail(CTA_EINVAL, "no forest %d", forest_id);
    *forest = it->second.forest;
  });
}

int cta_server_drop_forest(cta_server *s, int32_t forest_id) {
  return sguard([&] {
    if (!s) sfail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->forests.find(forest_id);
    if (it == s->forests.end()) sfail(CTA_EINVAL, "no forest %d", forest_id);
    it->second.dropped = true;
    s->release_forest_locked(forest_id);
  });
}

int cta_server_submit_predict(cta_server *s, int32_t forest_id, const float *X, int64_t R, int32_t F,
                              uint32_t *labels, uint64_t *ticket) {
  return cta_server_submit_predict_many(s, forest_id, 1, &X, &R, F, &labels, ticket);
}

int cta_server_submit_predict_many(cta_server *s, int32_t forest_id, int32_t count, const float *const *X,
                                   const int64_t *R, int32_t F, uint32_t *const *labels, uint64_t *tickets) {
  return sguard([&] {
    if (!s || !X || !R || !labels || !tickets) sfail(CTA_EINVAL, "NULL argument");
    if (count < 0) sfail(CTA_EINVAL, "count must be >= 0");
    const int64_t max_rows = s->o.max_batch_rows > 0 ? s->o.max_batch_rows : (1 << 20);
    for (int32_t i = 0; i < count; ++i) {
      if (!X[i] || !labels[i]) sfail(CTA_EINVAL, "NULL request %d", i);
      if (R[i] <= 0) sfail(CTA_EINVAL, "request %d: R must be positive", i);
      if (R[i] > max_rows) sfail(CTA_EINVAL, "request %d of %lld rows exceeds max_batch_rows", i, (long long)R[i]);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->forests.find(forest_id);
    if (it == s->forests.end() || it->second.dropped) sfail(CTA_EINVAL, "no forest %d", forest_id);
    if (F != it->second.F) sfail(CTA_EINVAL, "rows have %d features, the forest %d", F, it->second.F);
    for (int32_t i = 0; i < count; ++i) {
      const uint64_t t = s->new_ticket_locked();
      s->pq[s->lane_of(R[i])].push_back(PredReq{t, forest_id, X[i], R[i], labels[i]});
      tickets[i] = t;
    }
    s->cv_pred.notify_all();
  });
}

int cta_server_submit_fit(cta_server *s, const float *X, const uint32_t *y, int32_t N, int32_t F, int32_t K,
                          int32_t T, const cta_fit_opts *opts, uint64_t *ticket) {
  return sguard([&] {
    if (!s || !X || !y || !opts || !ticket) sfail(CTA_EINVAL, "NULL argument");
    if (N <= 0 || F <= 0 || K <= 0 || T <= 0) sfail(CTA_EINVAL, "sizes must be positive");
    FitJob j;
    j.X.assign(X, X + (size_t)N * F);
    j.y.assign(y, y + N);
    j.N = N;
    j.F = F;
    j.K = K;
    j.T = T;
    j.opts = *opts;
    uint64_t dsb = 0;
    fit_estimate(N, F, K, T, j.opts, &dsb, &j.ws_cap);
    j.est = dsb + j.ws_cap;
    std::lock_guard<std::mutex> lk(s->mu);
    j.ticket = s->new_ticket_locked();
    *ticket = j.ticket;
    s->fq.push_back(std::move(j));
    s->cv_fit.notify_all();
  });
}

int cta_server_wait(cta_server *s, uint64_t ticket, int64_t timeout_us, int32_t *status, int32_t *forest_id) {
  return sguard([&] {
    if (!s || !status) sfail(CTA_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) sfail(CTA_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    auto done = [&] { return s->tickets[ticket].status >= 0; };
    if (timeout_us < 0) s->cv_done.wait(lk, done);
    else s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), done);
    TicketState st = s->tickets[ticket];
    *status = st.status < 0 ? -1 : st.status;
    if (forest_id) *forest_id = st.forest;
    if (st.status >= 0) s->tickets.erase(ticket);
  });
}

int cta_server_wait_many(cta_server *s, int32_t count, const uint64_t *tickets, int64_t timeout_us, int32_t *status,
                         int32_t *forest_ids) {
  return sguard([&] {
    if (!s || count < 0 || (count > 0 && (!tickets || !status))) sfail(CTA_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    for (int32_t i = 0; i < count; ++i)
      if (s->tickets.find(tickets[i]) == s->tickets.end())
        sfail(CTA_EINVAL, "unknown ticket %llu", (unsigned long long)tickets[i]);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::microseconds(std::max<int64_t>(0, timeout_us));
    for (int32_t i = 0; i < count; ++i) {
      const uint64_t t = tickets[i];
      auto done = [&] { return s->tickets[t].status >= 0; };
      if (timeout_us < 0) s->cv_done.wait(lk, done);
      else s->cv_done.wait_until(lk, deadline, done);
      const TicketState st = s->tickets[t];
      status[i] = st.status < 0 ? -1 : st.status;
      if (forest_ids) forest_ids[i] = st.forest;
      if (st.status >= 0) s->tickets.erase(t);
    }
  });
}

int cta_server_get_stats(cta_server *s, cta_server_stats *out) {
  return sguard([&] {
    if (!s || !out) sfail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->predict_requests = s->pred_requests;
    out->predict_rows = s->pred_rows;
    out->predict_launches = s->pred_launches;
    out->fit_jobs = s->fit_done;
    out->fit_rejected = s->fit_rejected;
    out->hbm_budget = s->budget;
    out->hbm_in_use = s->in_use;
    out->hbm_peak = s->peak;
    out->resident_forests = (int32_t)s->forests.size();
    out->queued_predict = (int32_t)(s->pq[0].size() + s->pq[1].size());
    out->queued_fit = (int32_t)s->fq.size();
    if (!s->lat_ms.empty()) {
      std::vector<double> v = s->lat_ms;
      std::sort(v.begin(), v.end());
      auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
      out->p50_ms = pct(0.50);
      out->p99_ms = pct(0.99);
      double sum = 0;
      for (double x : v) sum += x;
      out->mean_ms = sum / v.size();
    }
    out->fit_mem_measured_max = s->fit_mem_max;
    out->fit_mem_ratio_max = s->fit_mem_ratio;
    const double span = s->t_last - s->t_first;
    out->rows_per_s = span > 0 ? (double)s->pred_rows / (span * 1e-3) : 0.0;
  });
}

int cta_server_reset_stats(cta_server *s) {
  return sguard([&] {
    if (!s) sfail(CTA_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->pred_requests = s->pred_rows = s->pred_launches = s->fit_done = s->fit_rejected = 0;
    s->fit_mem_max = 0;
    s->fit_mem_ratio = 0;
    s->t_first = s->t_last = 0;
    s->peak = s->in_use;
  });
}

}  // extern "C"

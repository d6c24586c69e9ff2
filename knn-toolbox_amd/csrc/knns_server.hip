// Serving loop: continuous batching of independent requests over a resident corpus.
//
// A request is (n queries, k). The worker thread takes every pending request in FIFO order that
// fits the largest bucket, packs their queries into consecutive slots, and replays the captured
// graph of the smallest bucket that holds them. Buckets are query-slot capacities 64 * 4^j up to
// max_batch; all share one device workspace (one batch runs at a time), sized for the largest.
// The graph is the engine's search sequence (enqueue_search) captured once per bucket with the
// bucket's slot count; the real query count, per-query k, shortlists and output offsets are
// copied to device buffers before each replay, and slots past the real count are inert
// (threshold -inf: they never append; init, merge and final skip them).
//
// Per-request results equal a search of the request alone: the key of a (query, row) pair does
// not depend on the other queries of the batch, and every query keeps the exact top-K' of its
// keys with the (key, index) order before the exact rerank.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "knns_engine.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  unsigned long long ticket = 0;
  std::vector<char> q;   // the request's queries (host copy)
  long n = 0;
  int k = 0;
  void* dist = nullptr;
  int32_t* idx = nullptr;
  knns_status status = KNNS_OK;
  std::string err;
  bool done = false;
  Clock::time_point t_submit, t_done;
};

struct Bucket {
  long cap = 0;
  knns::Plan P;
  hipGraphExec_t exec = nullptr;
};

}  // namespace

struct knns_server_s {
  knns_corpus c = nullptr;
  knns_server_config cfg{};
  int device = 0;
  int prec = KNNS_PREC_FP32;
  size_t esz = 4;       // bytes per query value (dtype of the corpus)
  size_t osz = 4;       // bytes per output distance
  int cmax = 1;
  hipStream_t stream = nullptr;
  std::vector<Bucket> buckets;
  knns::SearchBuffers B;
  std::vector<void*> dev_allocs;
  void* qstage = nullptr;   // device query rows of the batch
  int* n_dev = nullptr;     // device: real query count of the batch
  long workspace_bytes = 0;
  // pinned host staging
  char* h_q = nullptr;
  int* h_kq = nullptr;
  int* h_kc = nullptr;
  long* h_off = nullptr;
  int* h_n = nullptr;
  int* h_flags = nullptr;
  char* h_dist = nullptr;
  int32_t* h_idx = nullptr;
  std::vector<void*> host_allocs;
  // queue
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<unsigned long long, std::shared_ptr<Request>> live;
  unsigned long long next_ticket = 1;
  bool stop = false;
  std::thread worker;
  // statistics
  std::vector<double> lat_ms;
  long n_requests = 0, n_queries = 0, n_batches = 0;
  double busy_s = 0.0;

  int shortlist_of(int k) const {
    int cl = knns::default_shortlist(prec, k);
    cl = (int)std::min<long>(cl, c->m);
    return std::max(std::min(cl, 2048), 1);
  }
  long max_cap() const { return buckets.empty() ? 0 : buckets.back().cap; }

  void run_batch(std::vector<std::shared_ptr<Request>>& batch, long total);
  void loop();
  void release();
};

void knns_server_s::run_batch(std::vector<std::shared_ptr<Request>>& batch, long total) {
  const Bucket* bk = nullptr;
  for (const Bucket& b : buckets)
    if (b.cap >= total) { bk = &b; break; }
  auto fail_all = [&](knns_status st, const std::string& msg) {
    for (auto& r : batch) { r->status = st; r->err = msg; }
  };
  if (!bk) { fail_all(KNNS_ERR_INVALID, "batch larger than the largest bucket"); return; }
  // pack queries and per-query control values
  const size_t row = (size_t)c->d * esz;
  long slot = 0, out = 0;
  for (auto& r : batch) {
    std::memcpy(h_q + (size_t)slot * row, r->q.data(), (size_t)r->n * row);
    for (long i = 0; i < r->n; ++i) {
      h_kq[slot + i] = r->k;
      h_kc[slot + i] = shortlist_of(r->k);
      h_off[slot + i] = out + i * (long)r->k;
    }
    slot += r->n;
    out += r->n * (long)r->k;
  }
  *h_n = (int)total;
  hipError_t e = hipSuccess;
#define SRV_TRY(x) \
  if (e == hipSuccess) e = (x)
  SRV_TRY(hipMemcpyAsync(qstage, h_q, (size_t)total * row, hipMemcpyHostToDevice, stream));
  SRV_TRY(hipMemcpyAsync(B.kq, h_kq, (size_t)total * 4, hipMemcpyHostToDevice, stream));
  SRV_TRY(hipMemcpyAsync(B.kc, h_kc, (size_t)total * 4, hipMemcpyHostToDevice, stream));
  SRV_TRY(hipMemcpyAsync(B.off, h_off, (size_t)total * 8, hipMemcpyHostToDevice, stream));
  SRV_TRY(hipMemcpyAsync(n_dev, h_n, 4, hipMemcpyHostToDevice, stream));
  SRV_TRY(hipGraphLaunch(bk->exec, stream));
  SRV_TRY(hipMemcpyAsync(h_flags, B.flags, 16 * 4, hipMemcpyDeviceToHost, stream));
  SRV_TRY(hipMemcpyAsync(h_dist, B.odist, (size_t)out * osz, hipMemcpyDeviceToHost, stream));
  SRV_TRY(hipMemcpyAsync(h_idx, B.oidx, (size_t)out * 4, hipMemcpyDeviceToHost, stream));
  SRV_TRY(hipStreamSynchronize(stream));
#undef SRV_TRY
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail_all(KNNS_ERR_DEVICE, std::string("batch failed: ") + hipGetErrorString(e));
    return;
  }
  if (h_flags[0] != 0) {
    fail_all(KNNS_ERR_OVERFLOW, "candidate pool overflow: too many exact ties");
    return;
  }
  out = 0;
  for (auto& r : batch) {
    const size_t cnt = (size_t)r->n * r->k;
    std::memcpy(r->dist, h_dist + out * osz, cnt * osz);
    std::memcpy(r->idx, h_idx + out, cnt * 4);
    out += (long)cnt;
    r->status = KNNS_OK;
  }
}

void knns_server_s::loop() {
  (void)hipSetDevice(device);
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    long total = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      if (pending.empty()) return;   // stop requested and nothing left
      const long cap = max_cap();
      while (!pending.empty() && total + pending.front()->n <= cap) {
        total += pending.front()->n;
        batch.push_back(pending.front());
        pending.pop_front();
      }
    }
    const auto t0 = Clock::now();
    run_batch(batch, total);
    const auto t1 = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu);
      busy_s += std::chrono::duration<double>(t1 - t0).count();
      ++n_batches;
      for (auto& r : batch) {
        r->done = true;
        r->t_done = t1;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(t1 - r->t_submit).count());
        ++n_requests;
        n_queries += r->n;
      }
    }
    cv_done.notify_all();
  }
}

void knns_server_s::release() {
  for (Bucket& b : buckets)
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
  buckets.clear();
  for (void* p : dev_allocs) (void)hipFree(p);
  dev_allocs.clear();
  for (void* p : host_allocs) (void)hipHostFree(p);
  host_allocs.clear();
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

using knns::fail;

extern "C" {

knns_status knns_server_create(knns_server* out, knns_corpus c, const knns_server_config* cfg) {
  knns::g_err.clear();
  if (!out) return fail(KNNS_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (!c || !cfg) return fail(KNNS_ERR_INVALID, "null corpus or config");
  if (cfg->max_k < 1 || cfg->max_k > 1024) return fail(KNNS_ERR_INVALID, "max_k=%d outside [1, 1024]", cfg->max_k);
  if (cfg->max_batch < 1 || cfg->max_batch > (1L << 24))
    return fail(KNNS_ERR_INVALID, "max_batch=%ld outside [1, 2^24]", cfg->max_batch);
  if (cfg->mem_budget < 0) return fail(KNNS_ERR_INVALID, "negative memory budget");
  const int prec = cfg->precision;
  if (prec != KNNS_PREC_FP32 && prec != KNNS_PREC_BF16 && prec != KNNS_PREC_FP8 && prec != KNNS_PREC_FP64)
    return fail(KNNS_ERR_INVALID, "unknown precision %d", prec);
  if (prec == KNNS_PREC_FP64 && c->dtype != KNNS_F64) return fail(KNNS_ERR_INVALID, "FP64 precision needs a double corpus");
  if (prec == KNNS_PREC_BF16 && !c->x16) return fail(KNNS_ERR_INVALID, "BF16 precision needs a corpus built with KNNS_CORPUS_BF16");
  if (prec == KNNS_PREC_FP8 && !c->x8) return fail(KNNS_ERR_INVALID, "FP8 precision needs a corpus built with KNNS_CORPUS_FP8");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(KNNS_ERR_DEVICE, "no HIP device available");
  }
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(KNNS_ERR_DEVICE, "device query failed");
  const int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;

  auto S = std::make_unique<knns_server_s>();
  S->c = c;
  S->cfg = *cfg;
  S->device = dev;
  S->prec = prec;
  S->esz = c->dtype == KNNS_F64 ? 8 : 4;
  S->osz = c->dtype == KNNS_F64 ? 8 : 4;
  S->cmax = S->shortlist_of(cfg->max_k);
  // bucket capacities 64 * 4^j (and max_batch itself)
  std::vector<long> caps;
  for (long cap = 64; cap < cfg->max_batch; cap *= 4) caps.push_back(cap);
  caps.push_back(cfg->max_batch);
  // workspace: the maximum of each buffer over the buckets that fit the budget
  knns::SearchSizes z{};
  long budget = cfg->mem_budget;
  std::vector<knns::Plan> plans;
  for (long cap : caps) {
    const knns::Tuning tu = knns::tuning_for(prec, c->d, cfg->max_k, cap);
    knns::Plan P = knns::make_plan(c->m, c->d, cap, cfg->max_k, S->cmax, prec, cus, tu.ratio, 0, tu.min_waves);
    const knns::SearchSizes y = knns::search_sizes(c, P, cap, cap * (long)cfg->max_k, false, S->osz == 8);
    knns::SearchSizes t = z;
#define SRV_MAX(f) t.f = std::max(t.f, y.f)
    SRV_MAX(q32); SRV_MAX(q16); SRV_MAX(q8); SRV_MAX(q64); SRV_MAX(thr); SRV_MAX(kc); SRV_MAX(kq); SRV_MAX(off);
    SRV_MAX(sub); SRV_MAX(subcnt); SRV_MAX(carry); SRV_MAX(ccnt); SRV_MAX(flags); SRV_MAX(odist); SRV_MAX(oidx);
#undef SRV_MAX
    const long bytes = (long)(t.q32 + t.q16 + t.q8 + t.q64 + t.thr + t.kc + t.kq + t.off + t.sub + t.subcnt +
                              2 * t.carry + 2 * t.ccnt + t.flags + t.odist + t.oidx) +
                       cap * (long)c->d * (long)S->esz + 256;
    if (budget > 0 && bytes > budget) break;   // admission by memory: larger buckets do not fit
    z = t;
    S->workspace_bytes = bytes;
    plans.push_back(P);
    Bucket b;
    b.cap = cap;
    S->buckets.push_back(b);
  }
  if (S->buckets.empty())
    return fail(KNNS_ERR_NOMEM, "memory budget %ld bytes is below the smallest bucket's workspace", budget);
  for (size_t i = 0; i < S->buckets.size(); ++i) S->buckets[i].P = plans[i];
  const long cap_max = S->max_cap();

  auto dalloc = [&](size_t bytes, void** p) {
    if (bytes == 0) { *p = nullptr; return true; }
    if (hipMalloc(p, bytes) != hipSuccess) { (void)hipGetLastError(); *p = nullptr; return false; }
    S->dev_allocs.push_back(*p);
    return true;
  };
  auto halloc = [&](size_t bytes, void** p) {
    if (hipHostMalloc(p, bytes + 64, hipHostMallocDefault) != hipSuccess) { (void)hipGetLastError(); *p = nullptr; return false; }
    S->host_allocs.push_back(*p);
    return true;
  };
  knns::SearchBuffers& B = S->B;
  bool ok = dalloc(z.q32, &B.q32) && dalloc(z.q16, &B.q16) && dalloc(z.q8, &B.q8) && dalloc(z.q64, &B.q64) && dalloc(z.thr, &B.thr) &&
            dalloc(z.kc, &B.kc) && dalloc(z.kq, &B.kq) && dalloc(z.off, &B.off) && dalloc(z.sub, &B.sub) &&
            dalloc(z.subcnt, &B.subcnt) && dalloc(z.carry, &B.carry[0]) && dalloc(z.carry, &B.carry[1]) &&
            dalloc(z.ccnt, &B.ccnt[0]) && dalloc(z.ccnt, &B.ccnt[1]) && dalloc(z.flags, &B.flags) &&
            dalloc(z.odist, &B.odist) && dalloc(z.oidx, &B.oidx) &&
            dalloc((size_t)cap_max * c->d * S->esz, &S->qstage) && dalloc(256, (void**)&S->n_dev);
  void *hq, *hkq, *hkc, *hoff, *hn, *hfl, *hd, *hi;
  ok = ok && halloc((size_t)cap_max * c->d * S->esz, &hq) && halloc((size_t)cap_max * 4, &hkq) &&
       halloc((size_t)cap_max * 4, &hkc) && halloc((size_t)cap_max * 8, &hoff) && halloc(64, &hn) &&
       halloc(64, &hfl) && halloc(z.odist, &hd) && halloc(z.oidx, &hi);
  if (!ok) {
    S->release();
    return fail(KNNS_ERR_NOMEM, "server workspace allocation failed (%ld bytes)", S->workspace_bytes);
  }
  S->h_q = (char*)hq; S->h_kq = (int*)hkq; S->h_kc = (int*)hkc; S->h_off = (long*)hoff;
  S->h_n = (int*)hn; S->h_flags = (int*)hfl; S->h_dist = (char*)hd; S->h_idx = (int32_t*)hi;
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
    S->release();
    return fail(KNNS_ERR_DEVICE, "stream creation failed");
  }
  // one graph per bucket
  for (Bucket& b : S->buckets) {
    knns::SearchLaunch L;
    L.c = c;
    L.P = &b.P;
    L.n = b.cap;
    L.dtype = c->dtype;
    L.qsrc = S->qstage;
    L.uniform = false;
    L.kuni = 1;
    L.cmax = S->cmax;
    L.out64 = S->osz == 8;
    L.n_dev = S->n_dev;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t e2 = knns::enqueue_search(L, B, S->stream, nullptr);
      e = hipStreamEndCapture(S->stream, &g);
      if (e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      S->release();
      return fail(KNNS_ERR_DEVICE, "graph capture of bucket %ld failed: %s", b.cap, hipGetErrorString(e));
    }
  }
  knns_server s = S.release();
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return KNNS_OK;
}

knns_status knns_server_submit(knns_server s, const void* queries, int dtype, long n, int k,
                               void* dist, int32_t* idx, unsigned long long* ticket) {
  knns::g_err.clear();
  if (!s || !ticket) return fail(KNNS_ERR_INVALID, "null server or ticket");
  *ticket = 0;
  if (!queries || !dist || !idx) return fail(KNNS_ERR_INVALID, "null query/output pointer");
  if (dtype != s->c->dtype) return fail(KNNS_ERR_INVALID, "query dtype %d does not match corpus dtype %d", dtype, s->c->dtype);
  if (n < 1 || n > s->max_cap())
    return fail(KNNS_ERR_INVALID, "request of %ld queries outside [1, %ld] (the server's batch capacity)", n, s->max_cap());
  if (k < 1 || k > s->cfg.max_k) return fail(KNNS_ERR_INVALID, "k=%d outside [1, %d]", k, s->cfg.max_k);
  auto r = std::make_shared<Request>();
  r->n = n;
  r->k = k;
  r->dist = dist;
  r->idx = idx;
  r->q.assign((const char*)queries, (const char*)queries + (size_t)n * s->c->d * s->esz);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(KNNS_ERR_INVALID, "server is shutting down");
    r->ticket = s->next_ticket++;
    s->live[r->ticket] = r;
    s->pending.push_back(r);
  }
  s->cv_work.notify_one();
  *ticket = r->ticket;
  return KNNS_OK;
}

knns_status knns_server_wait(knns_server s, unsigned long long ticket, double timeout_s) {
  knns::g_err.clear();
  if (!s) return fail(KNNS_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return fail(KNNS_ERR_INVALID, "unknown ticket %llu", ticket);
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->done; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return fail(KNNS_ERR_TIMEOUT, "request %llu not done after %.3f s", ticket, timeout_s);
  }
  s->live.erase(ticket);
  if (r->status != KNNS_OK) return fail(r->status, "%s", r->err.c_str());
  return KNNS_OK;
}

knns_status knns_server_stats_get(knns_server s, knns_server_stats* out, int reset) {
  if (!s || !out) return fail(KNNS_ERR_INVALID, "null server or output");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->requests = s->n_requests;
  out->queries = s->n_queries;
  out->batches = s->n_batches;
  out->busy_s = s->busy_s;
  out->mean_batch = s->n_batches ? (double)s->n_queries / s->n_batches : 0.0;
  out->buckets = (long)s->buckets.size();
  out->max_batch = s->max_cap();
  out->workspace_bytes = s->workspace_bytes;
  if (!s->lat_ms.empty()) {
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = v.back();
  }
  if (reset) {
    s->lat_ms.clear();
    s->n_requests = s->n_queries = s->n_batches = 0;
    s->busy_s = 0.0;
  }
  return KNNS_OK;
}

knns_status knns_server_destroy(knns_server s) {
  if (!s) return KNNS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  s->release();
  delete s;
  return KNNS_OK;
}

}  // extern "C"

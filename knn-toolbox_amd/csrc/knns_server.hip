// Serving loop: continuous batching of independent requests over a resident corpus.
//
// A request is (n queries, k). Requests are split into pieces (ranges of their queries) that are
// packed, in arrival order, into batches of query slots. A server owns L lanes (default 2); each
// lane has its own device workspace, its own stream, its own pinned staging and one captured HIP
// graph per batch-capacity bucket (64 * 4^j slots up to max_batch). Two threads drive them:
//
//   dispatcher  waits for pending pieces and a free lane, packs as many pieces as the largest
//               bucket holds (a request larger than the room left is split: its first part goes
//               into this batch, the rest stays at the head of the queue), copies queries and the
//               per-query k / shortlist / output offsets to the lane, replays the graph of the
//               smallest bucket that holds the batch and queues the copies of the results back --
//               without waiting: while batch i runs, batch i+1 is packed, uploaded and launched on
//               the other lane, so their kernels overlap on the device;
//   completer   retires launched batches in order: waits for a lane's end event, re-searches the
//               queries the shortlist certificate could not prove complete, copies each piece's
//               results to its request and frees the lane.
//
// Admission is by memory: the lanes' workspaces are sized for the largest bucket whose L copies fit
// the memory budget; work beyond what the free lanes hold waits in the queue (deferred, never
// refused), and a request larger than the largest bucket is served in several batches.
//
// Per-request results equal a search of the request alone: the key of a (query, row) pair does not
// depend on the other queries of the batch, every query keeps the exact top-K' of its keys with the
// (key, index) order, and the exact rerank and the certificate are per query.
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

// A lane's stream. HIP pools its hardware queues: once a process holds GPU_MAX_HW_QUEUES streams
// (4 by default: the null stream, the engine's streams and a caller's own count), a new stream
// shares the least-used queue, and streams on one queue run strictly in order, so two lanes
// could serialise. A stream created with a CU mask always gets a queue of its own; the mask
// here enables every CU, so the lane's kernels see the whole device. (Such a stream is not
// created non-blocking: it orders with work on the null stream, which the server never uses.)
bool lane_stream_create(hipStream_t* s, int cus) {
  std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xFFFFFFFFu);
  if (hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data()) == hipSuccess) return true;
  (void)hipGetLastError();
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking) == hipSuccess;
}

struct Request {
  unsigned long long ticket = 0;
  std::vector<char> q;   // the request's queries (host copy)
  long n = 0;
  int k = 0;
  void* dist = nullptr;
  int32_t* idx = nullptr;
  long dispatched = 0;   // queries handed to batches so far
  long completed = 0;    // queries whose results are written
  knns_status status = KNNS_OK;
  std::string err;
  bool done = false;
  Clock::time_point t_submit, t_done;
};

// queries [q0, q0 + n) of a request, placed at batch slots [slot, slot + n)
struct Piece {
  std::shared_ptr<Request> r;
  long q0 = 0, n = 0, slot = 0, out = 0;
};

struct Bucket {
  long cap = 0;
  knns::Plan P;
};

struct Lane {
  hipStream_t stream = nullptr;
  knns::SearchBuffers B;
  void* qstage = nullptr;   // device query rows of the batch
  int* n_dev = nullptr;     // device: real query count of the batch
  std::vector<hipGraphExec_t> exec;   // one per bucket
  // pinned host staging
  char* h_q = nullptr;
  int* h_kq = nullptr;
  int* h_kc = nullptr;
  long* h_off = nullptr;
  int* h_n = nullptr;
  int* h_flags = nullptr;
  char* h_dist = nullptr;
  int32_t* h_idx = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  // the batch in flight
  std::vector<Piece> pieces;
  long total = 0, out_total = 0;
  bool busy = false;
  knns_status launch_status = KNNS_OK;
  std::string launch_err;
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
  std::vector<Bucket> buckets;
  std::vector<Lane> lanes;
  std::vector<void*> dev_allocs;
  std::vector<void*> host_allocs;
  long workspace_bytes = 0;
  hipEvent_t ev_ref = nullptr;   // time origin of the batch intervals
  // queue
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_inflight;
  std::deque<std::shared_ptr<Request>> pending;   // requests with queries not yet dispatched
  std::deque<int> inflight;                        // lanes launched, in launch order
  std::unordered_map<unsigned long long, std::shared_ptr<Request>> live;
  unsigned long long next_ticket = 1;
  bool stop = false;
  std::thread dispatcher, completer;
  // statistics
  std::vector<double> lat_ms;
  long n_requests = 0, n_queries = 0, n_batches = 0, n_overlapped = 0, n_researched = 0;
  double busy_s = 0.0, device_s = 0.0, last_end_ms = -1.0;
  Clock::time_point t_first_launch;
  bool any_launch = false;

  // K' per k (index k = 1..max_k): the built-in rule plus the tuning table's offset for that k
  // on this corpus at max_batch queries, the shortlist a plain search of that k would use
  std::vector<int> sl_k;
  int shortlist_of(int k) const { return sl_k[k]; }
  long max_cap() const { return buckets.empty() ? 0 : buckets.back().cap; }

  void launch(Lane& L);
  void retire(Lane& L);
  void dispatch_loop();
  void complete_loop();
  void release();
};

// Packs L.pieces into the lane's staging and enqueues upload, graph replay and result download on
// the lane's stream (no host synchronisation).
void knns_server_s::launch(Lane& L) {
  const Bucket* bk = nullptr;
  size_t bi = 0;
  for (; bi < buckets.size(); ++bi)
    if (buckets[bi].cap >= L.total) { bk = &buckets[bi]; break; }
  L.launch_status = KNNS_OK;
  if (!bk) {
    L.launch_status = KNNS_ERR_INVALID;
    L.launch_err = "batch larger than the largest bucket";
    return;
  }
  const size_t row = (size_t)c->d * esz;
  long out = 0;
  for (Piece& pc : L.pieces) {
    std::memcpy(L.h_q + (size_t)pc.slot * row, pc.r->q.data() + (size_t)pc.q0 * row, (size_t)pc.n * row);
    pc.out = out;
    for (long i = 0; i < pc.n; ++i) {
      L.h_kq[pc.slot + i] = pc.r->k;
      L.h_kc[pc.slot + i] = shortlist_of(pc.r->k);
      L.h_off[pc.slot + i] = out + i * (long)pc.r->k;
    }
    out += pc.n * (long)pc.r->k;
  }
  L.out_total = out;
  *L.h_n = (int)L.total;
  hipError_t e = hipSuccess;
#define SRV_TRY(x) \
  if (e == hipSuccess) e = (x)
  SRV_TRY(hipEventRecord(L.ev_start, L.stream));
  SRV_TRY(hipMemcpyAsync(L.qstage, L.h_q, (size_t)L.total * row, hipMemcpyHostToDevice, L.stream));
  SRV_TRY(hipMemcpyAsync(L.B.kq, L.h_kq, (size_t)L.total * 4, hipMemcpyHostToDevice, L.stream));
  SRV_TRY(hipMemcpyAsync(L.B.kc, L.h_kc, (size_t)L.total * 4, hipMemcpyHostToDevice, L.stream));
  SRV_TRY(hipMemcpyAsync(L.B.off, L.h_off, (size_t)L.total * 8, hipMemcpyHostToDevice, L.stream));
  SRV_TRY(hipMemcpyAsync(L.n_dev, L.h_n, 4, hipMemcpyHostToDevice, L.stream));
  SRV_TRY(hipGraphLaunch(L.exec[bi], L.stream));
  SRV_TRY(hipMemcpyAsync(L.h_flags, L.B.flags, 16 * 4, hipMemcpyDeviceToHost, L.stream));
  SRV_TRY(hipMemcpyAsync(L.h_dist, L.B.odist, (size_t)out * osz, hipMemcpyDeviceToHost, L.stream));
  SRV_TRY(hipMemcpyAsync(L.h_idx, L.B.oidx, (size_t)out * 4, hipMemcpyDeviceToHost, L.stream));
  SRV_TRY(hipEventRecord(L.ev_end, L.stream));
#undef SRV_TRY
  if (e != hipSuccess) {
    (void)hipGetLastError();
    L.launch_status = KNNS_ERR_DEVICE;
    L.launch_err = std::string("batch launch failed: ") + hipGetErrorString(e);
  }
}

// Waits for the lane's batch, re-searches its uncertified queries and writes every piece's results.
void knns_server_s::retire(Lane& L) {
  knns_status st = L.launch_status;
  std::string err = L.launch_err;
  if (st == KNNS_OK && hipEventSynchronize(L.ev_end) != hipSuccess) {
    (void)hipGetLastError();
    st = KNNS_ERR_DEVICE;
    err = "batch failed on the device";
  }
  if (st == KNNS_OK && L.h_flags[0] != 0) {
    st = KNNS_ERR_OVERFLOW;
    err = "candidate pool overflow: too many exact ties";
  }
  long researched = 0;
  if (st == KNNS_OK && L.h_flags[2] > 0 && prec != KNNS_PREC_FP8) {
    // queries whose shortlist the certificate could not prove complete: searched again with every
    // row inside the bound admitted, results rewritten in the lane's output buffers
    knns::ResearchIn R;
    R.c = c;
    R.prec = prec;
    R.dtype = c->dtype;
    R.out64 = osz == 8;
    R.row_begin = 0;
    R.row_end = c->m;
    R.qsrc = L.qstage;
    R.odist = L.B.odist;
    R.oidx = (int32_t*)L.B.oidx;
    std::vector<uint32_t> need(L.total);
    if (hipMemcpy(need.data(), L.B.thr, (size_t)L.total * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipGetLastError();
      st = KNNS_ERR_DEVICE;
      err = "certificate read-back failed";
    }
    for (long i = 0; st == KNNS_OK && i < L.total; ++i) {
      if (!need[i]) continue;
      R.qids.push_back((int)i);
      R.thr.push_back(need[i]);
      R.kq.push_back(L.h_kq[i]);
      R.off.push_back(L.h_off[i]);
      R.self.push_back(-1);
    }
    if (st == KNNS_OK) {
      st = knns::research_on_engine(R, L.stream, &researched);
      if (st != KNNS_OK) err = knns::g_err;
    }
    if (st == KNNS_OK &&
        (hipMemcpyAsync(L.h_dist, L.B.odist, (size_t)L.out_total * osz, hipMemcpyDeviceToHost, L.stream) != hipSuccess ||
         hipMemcpyAsync(L.h_idx, L.B.oidx, (size_t)L.out_total * 4, hipMemcpyDeviceToHost, L.stream) != hipSuccess ||
         hipStreamSynchronize(L.stream) != hipSuccess)) {
      (void)hipGetLastError();
      st = KNNS_ERR_DEVICE;
      err = "re-searched results could not be copied";
    }
  }
  float t0 = 0.f, t1 = 0.f;
  const bool timed = st == KNNS_OK && hipEventElapsedTime(&t0, ev_ref, L.ev_start) == hipSuccess &&
                     hipEventElapsedTime(&t1, ev_ref, L.ev_end) == hipSuccess;
  (void)hipGetLastError();
  // the callers' buffers stay valid until their request is done, which only happens below
  if (st == KNNS_OK)
    for (Piece& pc : L.pieces) {
      Request& r = *pc.r;
      std::memcpy((char*)r.dist + (size_t)pc.q0 * r.k * osz, L.h_dist + (size_t)pc.out * osz, (size_t)pc.n * r.k * osz);
      std::memcpy(r.idx + (size_t)pc.q0 * r.k, L.h_idx + pc.out, (size_t)pc.n * r.k * 4);
    }
  const auto now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(mu);
    for (Piece& pc : L.pieces) {
      Request& r = *pc.r;
      if (st != KNNS_OK && r.status == KNNS_OK) {
        r.status = st;
        r.err = err;
      }
      r.completed += pc.n;
      if (r.completed == r.n && r.dispatched == r.n) {
        r.done = true;
        r.t_done = now;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r.t_submit).count());
        ++n_requests;
        n_queries += r.n;
      }
    }
    ++n_batches;
    n_researched += researched;
    if (timed) {
      device_s += (t1 - t0) * 1e-3;
      if (last_end_ms >= 0.0 && t0 < last_end_ms) ++n_overlapped;   // began before the previous ended
      last_end_ms = std::max(last_end_ms, (double)t1);
    }
    busy_s = std::chrono::duration<double>(now - t_first_launch).count();
    L.pieces.clear();
    L.busy = false;
  }
  cv_done.notify_all();
  cv_work.notify_all();
}

void knns_server_s::dispatch_loop() {
  (void)hipSetDevice(device);
  for (;;) {
    int li = -1;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        if (pending.empty()) return stop;
        for (const Lane& L : lanes)
          if (!L.busy) return true;
        return false;
      });
      if (pending.empty()) break;   // stop requested and nothing left to dispatch
      int nfree = 0;
      for (size_t i = 0; i < lanes.size(); ++i)
        if (!lanes[i].busy) {
          if (li < 0) li = (int)i;
          ++nfree;
        }
      Lane& L = lanes[li];
      // pack pending queries in arrival order up to the largest bucket, splitting the last request;
      // with several lanes free the pending work is shared between them, so they run together
      long pend = 0;
      for (const auto& r : pending) pend += r->n - r->dispatched;
      const long share = (nfree > 1 && pend >= 256) ? (pend + nfree - 1) / nfree : pend;
      const long cap = std::min(max_cap(), std::max<long>(1, share));
      L.total = 0;
      while (!pending.empty() && L.total < cap) {
        auto r = pending.front();
        const long take = std::min(r->n - r->dispatched, cap - L.total);
        Piece pc;
        pc.r = r;
        pc.q0 = r->dispatched;
        pc.n = take;
        pc.slot = L.total;
        L.pieces.push_back(pc);
        r->dispatched += take;
        L.total += take;
        if (r->dispatched == r->n) pending.pop_front();
      }
      L.busy = true;
      if (!any_launch) { any_launch = true; t_first_launch = Clock::now(); }
    }
    launch(lanes[li]);
    {
      std::lock_guard<std::mutex> lk(mu);
      inflight.push_back(li);
    }
    cv_inflight.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    inflight.push_back(-1);   // no more batches
  }
  cv_inflight.notify_all();
}

void knns_server_s::complete_loop() {
  (void)hipSetDevice(device);
  for (;;) {
    int li;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_inflight.wait(lk, [&] { return !inflight.empty(); });
      li = inflight.front();
      inflight.pop_front();
    }
    if (li < 0) return;
    retire(lanes[li]);
  }
}

void knns_server_s::release() {
  for (Lane& L : lanes) {
    for (hipGraphExec_t g : L.exec)
      if (g) (void)hipGraphExecDestroy(g);
    L.exec.clear();
    if (L.ev_start) (void)hipEventDestroy(L.ev_start);
    if (L.ev_end) (void)hipEventDestroy(L.ev_end);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L.stream = nullptr;
  }
  lanes.clear();
  buckets.clear();
  for (void* p : dev_allocs) (void)hipFree(p);
  dev_allocs.clear();
  for (void* p : host_allocs) (void)hipHostFree(p);
  host_allocs.clear();
  if (ev_ref) (void)hipEventDestroy(ev_ref);
  ev_ref = nullptr;
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
  // at most two lanes: a third gained 2.7% of saturation throughput, and one of two measured
  // three-lane processes hung after its last batch (cause not found), so it is not offered
  if (cfg->lanes < 0 || cfg->lanes > 2) return fail(KNNS_ERR_INVALID, "lanes=%d outside [0, 2]", cfg->lanes);
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
  const int nl = cfg->lanes > 0 ? cfg->lanes : 2;

  auto S = std::make_unique<knns_server_s>();
  S->c = c;
  S->cfg = *cfg;
  S->device = dev;
  S->prec = prec;
  S->esz = c->dtype == KNNS_F64 ? 8 : 4;
  S->osz = c->dtype == KNNS_F64 ? 8 : 4;
  S->sl_k.assign((size_t)cfg->max_k + 1, 1);
  S->cmax = 1;
  for (int k = 1; k <= cfg->max_k; ++k) {
    const int extra = knns::tuning_for(prec, c->d, k, cfg->max_batch, c->m).sl_extra;
    S->sl_k[k] = knns::shortlist_for(prec, k, 0, extra, c->m);
    S->cmax = std::max(S->cmax, S->sl_k[k]);
  }
  // bucket capacities 64 * 4^j (and max_batch itself)
  std::vector<long> caps;
  for (long cap = 64; cap < cfg->max_batch; cap *= 4) caps.push_back(cap);
  caps.push_back(cfg->max_batch);
  // per-lane workspace: the maximum of each buffer over the buckets whose nl copies fit the budget
  knns::SearchSizes z{};
  long lane_bytes = 0;
  const long budget = cfg->mem_budget;
  for (long cap : caps) {
    const knns::Tuning tu = knns::tuning_for(prec, c->d, cfg->max_k, cap, c->m);
    knns::Plan P = knns::make_plan(c->m, c->d, cap, cfg->max_k, S->cmax, prec, cus, tu.ratio, 0, tu.min_waves,
                                   true, tu.qb, prec != KNNS_PREC_FP8);
    const knns::SearchSizes y = knns::search_sizes(c, P, cap, cap * (long)cfg->max_k, false, S->osz == 8);
    knns::SearchSizes t = z;
#define SRV_MAX(f) t.f = std::max(t.f, y.f)
    SRV_MAX(q32); SRV_MAX(q16); SRV_MAX(q8); SRV_MAX(q64); SRV_MAX(thr); SRV_MAX(kc); SRV_MAX(kq); SRV_MAX(off);
    SRV_MAX(sub); SRV_MAX(subcnt); SRV_MAX(carry); SRV_MAX(ccnt); SRV_MAX(flags); SRV_MAX(odist); SRV_MAX(oidx);
    SRV_MAX(qn2); SRV_MAX(qrho);
#undef SRV_MAX
    const long bytes = (long)(t.q32 + t.q16 + t.q8 + t.q64 + t.thr + t.kc + t.kq + t.off + t.sub + t.subcnt +
                              2 * t.carry + 2 * t.ccnt + t.flags + t.odist + t.oidx + t.qn2 + t.qrho) +
                       cap * (long)c->d * (long)S->esz + 256;
    if (budget > 0 && bytes * nl > budget) break;   // admission by memory: larger buckets do not fit
    z = t;
    lane_bytes = bytes;
    Bucket b;
    b.cap = cap;
    b.P = P;
    S->buckets.push_back(b);
  }
  if (S->buckets.empty())
    return fail(KNNS_ERR_NOMEM, "memory budget %ld bytes is below %d copies of the smallest bucket's workspace", budget, nl);
  S->workspace_bytes = lane_bytes * nl;
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
  S->lanes.resize(nl);
  bool ok = hipEventCreate(&S->ev_ref) == hipSuccess;
  for (Lane& L : S->lanes) {
    knns::SearchBuffers& B = L.B;
    ok = ok && dalloc(z.q32, &B.q32) && dalloc(z.q16, &B.q16) && dalloc(z.q8, &B.q8) && dalloc(z.q64, &B.q64) &&
         dalloc(z.thr, &B.thr) && dalloc(z.kc, &B.kc) && dalloc(z.kq, &B.kq) && dalloc(z.off, &B.off) &&
         dalloc(z.sub, &B.sub) && dalloc(z.subcnt, &B.subcnt) && dalloc(z.carry, &B.carry[0]) &&
         dalloc(z.carry, &B.carry[1]) && dalloc(z.ccnt, &B.ccnt[0]) && dalloc(z.ccnt, &B.ccnt[1]) &&
         dalloc(z.flags, &B.flags) && dalloc(z.odist, &B.odist) && dalloc(z.oidx, &B.oidx) &&
         dalloc(z.qn2, &B.qn2) && dalloc(z.qrho, &B.qrho) &&
         dalloc((size_t)cap_max * c->d * S->esz, &L.qstage) && dalloc(256, (void**)&L.n_dev);
    void *hq, *hkq, *hkc, *hoff, *hn, *hfl, *hd, *hi;
    ok = ok && halloc((size_t)cap_max * c->d * S->esz, &hq) && halloc((size_t)cap_max * 4, &hkq) &&
         halloc((size_t)cap_max * 4, &hkc) && halloc((size_t)cap_max * 8, &hoff) && halloc(64, &hn) &&
         halloc(64, &hfl) && halloc(z.odist, &hd) && halloc(z.oidx, &hi);
    if (!ok) break;
    L.h_q = (char*)hq; L.h_kq = (int*)hkq; L.h_kc = (int*)hkc; L.h_off = (long*)hoff;
    L.h_n = (int*)hn; L.h_flags = (int*)hfl; L.h_dist = (char*)hd; L.h_idx = (int32_t*)hi;
    ok = lane_stream_create(&L.stream, cus) &&
         hipEventCreate(&L.ev_start) == hipSuccess && hipEventCreate(&L.ev_end) == hipSuccess;
  }
  if (!ok) {
    S->release();
    return fail(KNNS_ERR_NOMEM, "server workspace allocation failed (%ld bytes)", S->workspace_bytes);
  }
  if (hipEventRecord(S->ev_ref, S->lanes[0].stream) != hipSuccess || hipStreamSynchronize(S->lanes[0].stream) != hipSuccess) {
    S->release();
    return fail(KNNS_ERR_DEVICE, "event set-up failed");
  }
  // one graph per (lane, bucket): the engine's search sequence over the lane's buffers
  for (Lane& L : S->lanes) {
    for (Bucket& b : S->buckets) {
      knns::SearchLaunch SL;
      SL.c = c;
      SL.P = &b.P;
      SL.n = b.cap;
      SL.dtype = c->dtype;
      SL.qsrc = L.qstage;
      SL.uniform = false;
      SL.kuni = 1;
      SL.cmax = S->cmax;
      SL.out64 = S->osz == 8;
      SL.n_dev = L.n_dev;
      SL.certify = true;
      SL.band = prec != KNNS_PREC_FP8;
      hipGraph_t g = nullptr;
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t e2 = knns::enqueue_search(SL, L.B, L.stream, nullptr);
        e = hipStreamEndCapture(L.stream, &g);
        if (e2 != hipSuccess) e = e2;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      L.exec.push_back(ex);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        S->release();
        return fail(KNNS_ERR_DEVICE, "graph capture of bucket %ld failed: %s", b.cap, hipGetErrorString(e));
      }
    }
  }
  knns_server s = S.release();
  s->completer = std::thread([s] { s->complete_loop(); });
  s->dispatcher = std::thread([s] { s->dispatch_loop(); });
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
  if (n < 1 || n > (1L << 30)) return fail(KNNS_ERR_INVALID, "request of %ld queries outside [1, 2^30]", n);
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
  s->cv_work.notify_all();
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
  out->lanes = (long)s->lanes.size();
  out->overlapped = s->n_overlapped;
  out->device_s = s->device_s;
  out->researched = s->n_researched;
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
    s->n_requests = s->n_queries = s->n_batches = s->n_overlapped = s->n_researched = 0;
    s->busy_s = s->device_s = 0.0;
    s->any_launch = false;
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
  if (s->dispatcher.joinable()) s->dispatcher.join();
  if (s->completer.joinable()) s->completer.join();
  s->release();
  delete s;
  return KNNS_OK;
}

}  // extern "C"

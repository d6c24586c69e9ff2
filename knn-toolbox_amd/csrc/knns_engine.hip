// Host engine and native C ABI (include/knns_native.h).
//
// One search = pack queries -> R threshold rounds (scan kernel over a growing corpus prefix,
// then merge) -> final rerank. The rounds replace the reference's corpus chunk x stream loop and
// host merge (src/cuknnBitonic.cu:550-589, src/cuknnHeap.cu:509-548): instead of merging P
// partial top-k lists on the host, every round starts from the exact K-th candidate of all the
// corpus rows seen so far, so the filter inside the scan kernel stays tight.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include <dlfcn.h>
#include <string>
#include <vector>

#include "knns_engine.h"

namespace knns {

thread_local std::string g_err;

knns_status fail(knns_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

namespace {

#define KNNS_HIP(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(KNNS_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                      \
  } while (0)

enum Slot {
  W_Q32, W_Q16, W_Q8, W_Q64, W_STAGE, W_THR, W_KC, W_KQ, W_OFF, W_SELF, W_SUB, W_SUBCNT, W_CARRY0,
  W_CARRY1, W_CCNT0, W_CCNT1, W_FLAGS, W_ODIST, W_OIDX, W_QN2, W_QRHO, W_RQ, W_RCTL, W_NSLOTS
};

struct Engine {
  std::recursive_mutex mu;
  bool ready = false;
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  hipStream_t ustream = nullptr;   // corpus uploads that overlap a search (knns_search_once); created
                                   // at its first use: a process has few hardware queues (4 by
                                   // default) and streams beyond them share queues and serialise
  void* buf[W_NSLOTS] = {};
  size_t cap[W_NSLOTS] = {};
  std::vector<hipEvent_t> events;
  knns_stats last{};

  knns_status init() {
    if (ready) return KNNS_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
      return fail(KNNS_ERR_DEVICE, "no HIP device available");
    KNNS_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    KNNS_HIP(hipGetDeviceProperties(&prop, device));
    cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    KNNS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ready = true;
    return KNNS_OK;
  }

  knns_status get(Slot s, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    if (cap[s] < bytes) {
      if (buf[s]) (void)hipFree(buf[s]);
      buf[s] = nullptr;
      cap[s] = 0;
      const size_t want = bytes + bytes / 8;
      if (hipMalloc(&buf[s], want) != hipSuccess) {
        (void)hipGetLastError();
        buf[s] = nullptr;
        return fail(KNNS_ERR_NOMEM, "device allocation of %zu bytes failed", want);
      }
      cap[s] = want;
    }
    *out = buf[s];
    return KNNS_OK;
  }

  void* pin = nullptr;
  size_t pin_cap = 0;
  // grow-only pinned host staging (reused only after the previous search synchronised)
  knns_status pinned(size_t bytes, void** out) {
    if (pin_cap < bytes) {
      if (pin) (void)hipHostFree(pin);
      pin = nullptr;
      pin_cap = 0;
      const size_t want = bytes + bytes / 4 + 4096;
      if (hipHostMalloc(&pin, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        pin = nullptr;
        return fail(KNNS_ERR_NOMEM, "pinned host allocation of %zu bytes failed", want);
      }
      pin_cap = want;
    }
    *out = pin;
    return KNNS_OK;
  }

  void* upin = nullptr;
  size_t upin_cap = 0;
  // grow-only pinned host buffer of the corpus upload (two halves, double-buffered)
  knns_status upload_pinned(size_t bytes, void** out) {
    if (upin_cap < bytes) {
      if (upin) (void)hipHostFree(upin);
      upin = nullptr;
      upin_cap = 0;
      if (hipHostMalloc(&upin, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        upin = nullptr;
        return fail(KNNS_ERR_NOMEM, "pinned host allocation of %zu bytes failed", bytes);
      }
      upin_cap = bytes;
    }
    *out = upin;
    return KNNS_OK;
  }

  void release() {
    for (int s = 0; s < W_NSLOTS; ++s) {
      if (buf[s]) (void)hipFree(buf[s]);
      buf[s] = nullptr;
      cap[s] = 0;
    }
  }

  hipEvent_t event(size_t i) {
    while (events.size() <= i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      events.push_back(e);
    }
    return events[i];
  }
};

Engine& engine() {
  static Engine* e = new Engine();  // never destroyed: HIP may already be torn down at exit
  return *e;
}

}  // namespace

// ---- tuning table ----
// Text file, one entry per line: precision d k n m min_waves ratio qb sl_extra [measured ms]
// (m = corpus rows, qb = query blocks per scan wave, 0 = built-in rule; sl_extra = shortlist
// offset from the precision's built-in rule); '#' starts a comment. Loaded at the first search from $KNNS_TUNING_TABLE, else from
// <directory of libknns_amd.so>/../tuning/gfx950.tsv; knns_set_tuning_table replaces it.
namespace {
struct TuneEntry {
  int prec, d, k;
  long n, m;
  Tuning t;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;
bool g_tune_loaded = false;

knns_status load_tuning(const char* path, std::vector<TuneEntry>* out) {
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(KNNS_ERR_INVALID, "cannot open tuning table %s", path);
  std::vector<TuneEntry> v;
  char line[512];
  int ln = 0;
  while (std::fgets(line, sizeof(line), f)) {
    ++ln;
    char* h = std::strchr(line, '#');
    if (h) *h = '\0';
    char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\0' || *p == '\n' || *p == '\r') continue;
    char prec[16];
    TuneEntry e{};
    if (std::sscanf(p, "%15s %d %d %ld %ld %d %d %d %d", prec, &e.d, &e.k, &e.n, &e.m, &e.t.min_waves, &e.t.ratio,
                    &e.t.qb, &e.t.sl_extra) != 9) {
      std::fclose(f);
      return fail(KNNS_ERR_INVALID, "%s:%d: expected 'precision d k n m min_waves ratio qb sl_extra'", path, ln);
    }
    if (!std::strcmp(prec, "fp32")) e.prec = KNNS_PREC_FP32;
    else if (!std::strcmp(prec, "bf16")) e.prec = KNNS_PREC_BF16;
    else if (!std::strcmp(prec, "fp8")) e.prec = KNNS_PREC_FP8;
    else if (!std::strcmp(prec, "fp64")) e.prec = KNNS_PREC_FP64;
    else {
      std::fclose(f);
      return fail(KNNS_ERR_INVALID, "%s:%d: unknown precision '%s'", path, ln, prec);
    }
    if (e.d < 1 || e.k < 1 || e.n < 1 || e.m < 1 || e.t.min_waves < 1 || e.t.min_waves > 64 ||
        (e.t.ratio != 0 && (e.t.ratio < 2 || e.t.ratio > 16)) || e.t.qb < 0 || e.t.qb > 4 ||
        e.t.sl_extra < -1024 || e.t.sl_extra > 1024) {
      std::fclose(f);
      return fail(KNNS_ERR_INVALID, "%s:%d: value out of range", path, ln);
    }
    v.push_back(e);
  }
  std::fclose(f);
  *out = std::move(v);
  return KNNS_OK;
}

std::string default_tuning_path() {
  if (const char* env = std::getenv("KNNS_TUNING_TABLE")) return env;
  Dl_info info;
  if (dladdr((void*)&default_tuning_path, &info) && info.dli_fname) {
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tuning/gfx950.tsv";
  }
  return "";
}
}  // namespace

Tuning tuning_for(int prec, int d, int k, long n, long m) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_loaded) {
    g_tune_loaded = true;
    const std::string path = default_tuning_path();
    std::vector<TuneEntry> v;
    const std::string keep = g_err;   // a missing default table is not an error of the caller
    if (!path.empty() && load_tuning(path.c_str(), &v) == KNNS_OK) g_tune = std::move(v);
    g_err = keep;
  }
  // FP64 searches run the FP32 scan and shortlist rule: without FP64 entries they take the FP32 ones
  int use = prec;
  if (prec == KNNS_PREC_FP64 &&
      std::none_of(g_tune.begin(), g_tune.end(), [](const TuneEntry& e) { return e.prec == KNNS_PREC_FP64; }))
    use = KNNS_PREC_FP32;
  Tuning best;
  double bd = 1e30;
  for (const TuneEntry& e : g_tune) {
    if (e.prec != use) continue;
    const double dist = std::fabs(std::log2((double)d / e.d)) + std::fabs(std::log2((double)k / e.k)) +
                        std::fabs(std::log2((double)n / e.n)) + std::fabs(std::log2((double)m / e.m));
    if (dist < bd) { bd = dist; best = e.t; }
  }
  return best;
}

// FP8 operand row: d bytes rounded up to a row width the FP8 scan kernel is built for (64, 128 or
// 256 bytes: its LDS swizzle needs a power-of-two count of 16-byte units); wider rows are rejected
int fp8_row_bytes(int d) { return d <= 64 ? 64 : (d <= 128 ? 128 : (int)round_up(d, 256)); }

// padded dimension and MFMA chunk for a dimension d
void dims_for(int d, int* dp, int* dc) {
  const int d16 = (int)round_up(d, 16);
  if (d16 <= 32) *dc = 32;
  else if (d16 <= 64) *dc = 64;
  else if (d16 <= 96) *dc = 96;
  else *dc = 128;
  *dp = d16 <= 128 ? *dc : (int)round_up(d, 128);
}

// Shortlist sizes K' per precision: the K'-th approximate key sets each threshold. Exactness does
// not depend on them: every threshold also keeps the certificate band above the K'-th key (rows
// whose distance the key error bound cannot separate from it, BandArgs), and the certificate
// re-searches any query whose final shortlist it cannot prove complete. A fixed margin instead of
// the band needs K' far beyond k on dense data (1M x 128 mixture, k = 100: K' = 296 for all but 16
// of 10k queries at BF16; 10M rows: none at 296). FP8, the approximate form, has no band: its K'
// sets its recall.
int default_shortlist(int prec, int k) {
  switch (prec) {
    case KNNS_PREC_BF16: return k + k / 2 + 16;
    case KNNS_PREC_FP8: return 4 * k + 32;
    default: return k + 16;
  }
}

int shortlist_for(int prec, int k, int requested, int sl_extra, long m_eff) {
  int cl = requested > 0 ? std::max(requested, k) : std::max(k, default_shortlist(prec, k) + sl_extra);
  cl = (int)std::min<long>(cl, m_eff);
  return std::max(std::min(cl, 2048), 1);
}

Plan make_plan(long m, int d, long n, int kmax, int cmax, int prec, int cus, int ratio_req,
               int slices_req, int min_waves, bool allow_dense, int qb_req, bool band) {
  Plan P{};
  dims_for(d, &P.dp, &P.dc);
  P.prec_kernel = prec == KNNS_PREC_BF16 ? PREC_BF16 : (prec == KNNS_PREC_FP8 ? PREC_FP8 : PREC_F32);
  if (P.prec_kernel == PREC_FP8) P.dc = fp8_row_bytes(d);
  // query blocks per scan wave: 2 for the low-precision scans; with enough queries to fill the
  // wider tiles, 3 or 4 for FP8 rows of up to 128 bytes (10M x 96, 10k queries: scan 11.85 ms at
  // 2 blocks, 10.92 at 3, 9.96 at 4) and 3 for 128-dim BF16 rows (1M x 128, 10k queries: -2%)
  P.qb = (P.prec_kernel != PREC_F32 && n > 32) ? 2 : 1;
  if (P.prec_kernel == PREC_FP8 && P.dc <= 128 && n >= 2048) P.qb = n >= 4096 ? 4 : 3;
  if (P.prec_kernel == PREC_BF16 && P.dp == 128 && P.dc == 128 && n >= 4096) P.qb = 3;
  if (qb_req > 0 && scan_qb_ok(P.prec_kernel, P.dp, P.dc, qb_req)) P.qb = qb_req;
  const int qw = 32 * P.qb;
  P.groups = (int)round_up((n + qw - 1) / qw, 4);  // a scan workgroup runs 4 query groups
  P.n_pad = P.groups * qw;
  P.kmax = kmax;
  P.cmax = cmax;
  // the band keeps rows beyond the K'-th: room for up to 4 K' (at most 1024, the rerank's wave form)
  P.carry_cap = band ? std::max(cmax, std::min(1024, 4 * cmax)) : cmax;
  // scan workgroups: 4 query groups each, 2 resident per CU (1 for the 3- and 4-block scans of
  // wide rows). Choose the slice count so that the workgroups of a large round fill whole waves of
  // resident workgroups (a last wave that is
  // mostly empty leaves the chip idle): the smallest S with >= min_waves waves and >= 95% of the
  // last wave used, else the S with the best fill. min_waves comes from the tuning table
  // (default 5: on 1M x 128 with 10k queries 5 waves beat 4 by 5%, more add merge sources).
  const long wg_rows = P.groups / 4;
  const long cap = (long)cus * scan_wgs_per_cu(P.prec_kernel, P.dp, P.dc, P.qb);
  int smax = 1;
  {
    double best = -1.0;
    for (int S = 1; S <= 128; ++S) {
      const long wgs = wg_rows * S;
      const long waves = (wgs + cap - 1) / cap;
      const double fill = (double)wgs / (double)(waves * cap);
      if (waves >= min_waves && fill >= 0.95) { smax = S; best = 2.0; break; }
      const double score = fill * std::min<double>(1.0, (double)wgs / ((double)min_waves * cap));
      if (score > best) { best = score; smax = S; }
    }
  }
  if (slices_req > 0) smax = std::min(128, slices_req);
  // growth factor of the corpus prefix between threshold rounds: a fixed ratio if requested,
  // else x4 while the prefix is under m/16 (cheap rounds, fewer merges) and x2 after (the large
  // rounds, where a tighter threshold saves filter work). Measured on 1M x 128 (k = 100, BF16):
  // x8 up to m/8 makes 5 rounds instead of 8 but costs 0.12 ms more per search (its large third
  // round admits ~290 candidates per query; the merges it saves take ~15 us each).
  const long gr = scan_stage_rows(P.prec_kernel, P.dp, P.dc);  // boundary granularity (rows)
  P.sub_cap = (int)round_up(std::max<long>(cmax + 96, 256), 64);
  // The first round has no threshold: every row of it is a candidate. It is written densely
  // (no filter) and made as large as the merge takes in registers (1024 entries per query), which
  // replaces the two smallest threshold rounds of a pure x4 schedule.
  long b = 0;
  long n0 = std::min<long>(m, round_up(std::max<long>(std::max<long>(2L * cmax, 1024), gr), gr));
  long e = n0;
  int smax_used = 1;
  while (b < m) {
    Round r;
    r.begin = b;
    r.end = e;
    const long rows = e - b;
    // small rounds: at least one stage per slice, as many slices as that allows (more waves
    // share the latency-bound appends of the first rounds)
    int s = (int)std::min<long>(smax, std::max<long>(1, rows / gr));
    r.slice_rows = round_up((rows + s - 1) / s, gr);
    r.slices = (int)((rows + r.slice_rows - 1) / r.slice_rows);
    r.dense = (allow_dense && b == 0 && r.slice_rows <= P.sub_cap) ? 1 : 0;
    smax_used = std::max(smax_used, r.slices);
    P.rounds.push_back(r);
    b = e;
    const long ratio = ratio_req >= 2 ? ratio_req : (e * 16 < m ? 4 : 2);
    e = std::min<long>(m, round_up(e * ratio, gr));
  }
  P.slices_max = smax_used;
  return P;
}

double key_gamma(const Plan& P) {
  // roundings along one key's accumulation: FP32 products are rounded (two per k-step of the
  // 32x32x2 MFMA, plus the hi/lo norm step); BF16 and FP8 products are exact in FP32 and each
  // k-step adds its products to the accumulator; the norm step adds up to 8 terms
  const double n = P.prec_kernel == PREC_F32 ? 2.0 * (P.dp + 4) : (double)(P.prec_kernel == PREC_FP8 ? P.dc : P.dp) + 16.0;
  const double u = 0x1p-23;
  // + the norm's split into 3 BF16 terms (relative 2^-27) or 2 FP32 terms
  return n * u / (1.0 - n * u) + 0x1p-26;
}

SearchSizes search_sizes(const knns_corpus_s* c, const Plan& P, long n, long total_out, bool excl,
                         bool out64) {
  const long n_pad = P.n_pad;
  SearchSizes z{};
  z.q32 = (size_t)n_pad * c->ld32 * 4;
  z.q16 = P.prec_kernel == PREC_BF16 ? (size_t)n_pad * c->ld16 * 2 : 0;
  z.q8 = P.prec_kernel == PREC_FP8 ? (size_t)n_pad * c->ld8 : 0;
  z.q64 = c->dtype == KNNS_F64 ? (size_t)n * c->d * 8 : 0;
  z.thr = (size_t)n_pad * 4;
  z.kc = (size_t)n_pad * 4;
  z.kq = (size_t)n * 4;
  z.off = (size_t)n * 8;
  z.self = excl ? (size_t)n_pad * 4 : 0;
  z.sub = (size_t)n_pad * P.slices_max * 2 * P.sub_cap * 8;
  z.subcnt = (size_t)n_pad * P.slices_max * 2 * 4;
  z.carry = (size_t)n * P.carry_cap * 8;
  z.ccnt = (size_t)n * 4;
  z.flags = 16 * 4;
  z.odist = (size_t)total_out * (out64 ? 8 : 4);
  z.oidx = (size_t)total_out * 4;
  z.qn2 = (size_t)n_pad * 8;
  z.qrho = (size_t)n_pad * 8;
  return z;
}

// memcpy split over up to 8 host threads (one thread moves about 30 GB/s on these hosts, the DMA
// from pinned memory about 57 GB/s)
void parallel_copy(void* dst, const void* src, size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t parts = std::min<size_t>(std::min(hw, 8u), bytes >> 22);   // >= 4 MB per thread
  if (parts <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t step = ((bytes + parts - 1) / parts + 4095) & ~(size_t)4095;
  std::vector<std::thread> th;
  th.reserve(parts);
  try {
    for (size_t o = step; o < bytes; o += step)
      th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, std::min(step, bytes - o)); });
  } catch (...) {   // no thread: the caller's thread copies what is left
    size_t o = step * (th.size() + 1);
    if (o < bytes) std::memcpy((char*)dst + o, (const char*)src + o, bytes - o);
  }
  std::memcpy(dst, src, std::min(step, bytes));
  for (auto& t : th) t.join();
}

// Chunked fill of a corpus' packed rows [0, rows_pad). Host data goes up in chunks of about
// 64 MB through the two halves of a pinned buffer: host threads fill one half while the DMA and the
// pack kernel of the previous chunk run on `st` (pageable copies reach about half the pinned
// rate). Device data is packed in one chunk. `ready` is recorded after each chunk's pack.
struct Upload {
  knns_corpus_s* c = nullptr;
  const void* data = nullptr;
  bool dev = false;
  hipStream_t st = nullptr;
  long row_bytes = 0, chunk = 0, r0 = 0;
  int ci = 0;
  void* stage = nullptr;
  char* hp[2] = {nullptr, nullptr};
  hipEvent_t freed[2] = {nullptr, nullptr};   // DMA out of pinned half h done
  bool pending[2] = {false, false};
  hipEvent_t ready = nullptr;

  knns_status start(knns_corpus_s* c_, const void* data_, bool dev_, hipStream_t st_) {
    c = c_;
    data = data_;
    dev = dev_;
    st = st_;
    row_bytes = (long)c->d * (c->dtype == KNNS_F64 ? 8L : 4L);
    chunk = dev ? c->rows_pad : std::min<long>(c->rows_pad, std::max<long>(1, (64L << 20) / row_bytes));
    if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) {
      ready = nullptr;
      return fail(KNNS_ERR_DEVICE, "hipEventCreate failed");
    }
    // the certificate maxima start at zero, before the first pack raises them
    if (launch_fill_u32(c->bnd, 0u, BND_SLOTS, st) != hipSuccess) return fail(KNNS_ERR_DEVICE, "bounds clear failed");
    if (dev) return KNNS_OK;
    if (hipMalloc(&stage, (size_t)chunk * row_bytes) != hipSuccess) {
      (void)hipGetLastError();
      stage = nullptr;
      return fail(KNNS_ERR_NOMEM, "corpus staging buffer");
    }
    void* pp = nullptr;
    if (knns_status ps = engine().upload_pinned(2 * (size_t)chunk * row_bytes, &pp)) return ps;
    hp[0] = (char*)pp;
    hp[1] = (char*)pp + (size_t)chunk * row_bytes;
    for (auto& e : freed)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        e = nullptr;
        return fail(KNNS_ERR_DEVICE, "hipEventCreate failed");
      }
    return KNNS_OK;
  }

  long rows_done() const { return r0; }

  // enqueue the next chunk (no-op once every row is enqueued)
  knns_status next() {
    if (r0 >= c->rows_pad) return KNNS_OK;
    const long rows_w = std::min(chunk, c->rows_pad - r0);
    const long rows = std::max<long>(0, std::min(rows_w, c->m - r0));
    const void* src = (const char*)data + (size_t)r0 * row_bytes;
    const int h = ci & 1;
    if (!dev && rows > 0) {
      if (pending[h] && hipEventSynchronize(freed[h]) != hipSuccess)
        return fail(KNNS_ERR_DEVICE, "corpus upload failed");
      parallel_copy(hp[h], src, (size_t)rows * row_bytes);
      // the DMA into `stage` follows the previous chunk's pack on the same stream
      if (hipMemcpyAsync(stage, hp[h], (size_t)rows * row_bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipEventRecord(freed[h], st) != hipSuccess)
        return fail(KNNS_ERR_DEVICE, "corpus upload failed");
      pending[h] = true;
      src = stage;
    }
    PackArgs pa{};
    pa.src = src;
    pa.src_f64 = c->dtype == KNNS_F64;
    pa.rows = rows;
    pa.rows_pad = rows_w;
    pa.d = c->d;
    pa.ld_src = c->d;
    pa.x32 = c->x32 + (size_t)r0 * c->ld32;
    pa.ld32 = c->ld32;
    pa.x16 = c->x16 ? c->x16 + (size_t)r0 * c->ld16 : nullptr;
    pa.ld16 = c->ld16;
    pa.x64 = c->x64 ? c->x64 + (size_t)r0 * c->d : nullptr;
    pa.dp = c->dp;
    pa.is_query = 0;
    pa.x8 = c->x8 ? c->x8 + (size_t)r0 * c->ld8 : nullptr;
    pa.ld8 = c->ld8;
    pa.dp8 = c->dp8;
    pa.bnd = c->bnd;
    pa.rrho = c->rrho ? c->rrho + r0 : nullptr;
    pa.rrho_ld = c->rows_pad;
    if (launch_pack(pa, st) != hipSuccess || (c->x8 && launch_pack_fp8(pa, st) != hipSuccess))
      return fail(KNNS_ERR_DEVICE, "corpus pack launch failed");
    if (hipEventRecord(ready, st) != hipSuccess) return fail(KNNS_ERR_DEVICE, "hipEventRecord failed");
    r0 += rows_w;
    ++ci;
    return KNNS_OK;
  }

  // wait for everything enqueued, release the staging (idempotent)
  knns_status finish() {
    knns_status s = KNNS_OK;
    if (st && hipStreamSynchronize(st) != hipSuccess) s = fail(KNNS_ERR_DEVICE, "corpus pack failed");
    st = nullptr;
    if (stage) (void)hipFree(stage);
    stage = nullptr;
    for (auto& e : freed)
      if (e) (void)hipEventDestroy(e);
    freed[0] = freed[1] = nullptr;
    if (ready) (void)hipEventDestroy(ready);
    ready = nullptr;
    return s;
  }

  ~Upload() { (void)finish(); }
};

// Enqueue the upload until `rows` rows (capped at rows_pad) are packed; `st` waits for them.
hipError_t feed_rows(Upload& u, long rows, hipStream_t st) {
  rows = std::min(rows, u.c->rows_pad);
  while (u.rows_done() < rows)
    if (u.next() != KNNS_OK) return hipErrorUnknown;
  return hipStreamWaitEvent(st, u.ready, 0);
}

hipError_t enqueue_search(const SearchLaunch& L, const SearchBuffers& B, hipStream_t st,
                          const LaunchHook* hook) {
  const knns_corpus_s* c = L.c;
  const Plan& P = *L.P;
  const long n = L.n, n_pad = P.n_pad, m = c->m;
  hipError_t e;
#define KNNS_TRY(x) \
  if ((e = (x)) != hipSuccess) return e
  // queries -> packed operand rows
  PackArgs pa{};
  pa.src = L.qsrc;
  pa.src_f64 = L.dtype == KNNS_F64;
  pa.rows = n;
  pa.rows_pad = n_pad;
  pa.d = c->d;
  pa.ld_src = c->d;
  pa.x32 = (float*)B.q32;
  pa.ld32 = c->ld32;
  pa.x16 = (uint16_t*)B.q16;
  pa.ld16 = c->ld16;
  pa.x64 = (double*)B.q64;
  pa.dp = c->dp;
  pa.is_query = 1;
  pa.qn2 = (double*)B.qn2;
  pa.qrho = (double*)B.qrho;
  pa.qprec = P.prec_kernel;
  KNNS_TRY(launch_pack(pa, st));
  if (P.prec_kernel == PREC_FP8) {
    pa.x8 = (uint8_t*)B.q8;
    pa.ld8 = c->ld8;
    pa.dp8 = c->dp8;
    KNNS_TRY(launch_pack_fp8(pa, st));
  }
  KNNS_TRY(launch_init_queries((uint32_t*)B.thr, (int*)B.kc, (int*)B.kq, (long*)B.off,
                               L.self_preset ? nullptr : (int*)B.self, (int*)B.ccnt[0], n, n_pad, L.uniform ? L.kuni : 1,
                               L.cmax, L.self_offset, m, L.n_dev, L.uniform ? 1 : 0, L.thr_preset ? 0 : 1, st));
  KNNS_TRY(launch_fill_i32((int*)B.flags, 0, 16, st));

  const int pk = P.prec_kernel;
  BandArgs band{};
  band.on = L.band ? 1 : 0;
  band.qn2 = (const double*)B.qn2;
  band.qrho = (const double*)B.qrho;
  band.bnd = c->bnd;
  band.bslot = pk;
  band.gamma = key_gamma(P);
  if (band.on && (!B.qn2 || !c->bnd)) return hipErrorInvalidValue;
  const void* xop = pk == PREC_BF16 ? (const void*)c->x16 : (pk == PREC_FP8 ? (const void*)c->x8 : (const void*)c->x32);
  const void* qop = pk == PREC_BF16 ? B.q16 : (pk == PREC_FP8 ? B.q8 : B.q32);
  const long ldx = pk == PREC_BF16 ? c->ld16 : (pk == PREC_FP8 ? c->ld8 : c->ld32);
  int cur = 0;
  for (size_t r = 0; r < P.rounds.size(); ++r) {
    const Round& R = P.rounds[r];
    ScanArgs sa{};
    sa.X = xop;
    sa.Qp = qop;
    sa.ldx = ldx;
    sa.ldq = ldx;
    sa.dp = pk == PREC_FP8 ? c->dp8 : P.dp;
    sa.row_begin = R.begin;
    sa.row_end = R.end;
    sa.slice_rows = R.slice_rows;
    sa.slices = R.slices;
    sa.slices_max = P.slices_max;
    sa.n_pad = (int)n_pad;
    sa.thr = (const uint32_t*)B.thr;
    sa.kc = (const int*)B.kc;
    sa.sub = (uint64_t*)B.sub;
    sa.sub_cnt = (int*)B.subcnt;
    sa.sub_cap = P.sub_cap;
    sa.self_ids = (const int*)B.self;
    sa.flags = (int*)B.flags;
    sa.dense = R.dense;
    sa.n_dev = L.n_dev;
    sa.band = band;
    sa.thr_w = (uint32_t*)B.thr;
    sa.wg_rows = P.groups / 4;
    // an overlapped upload: the scan reads up to one stage (<= 128 rows) past its round's end
    if (L.feed) KNNS_TRY(feed_rows(*L.feed, R.end + 128, st));
    if (hook) (*hook)(0);
    KNNS_TRY(launch_scan(P.prec_kernel, P.qb, P.dc, sa, P.groups, st));
    if (hook) (*hook)(1);

    MergeArgs ma{};
    ma.carry_in = (const uint64_t*)B.carry[cur];
    ma.carry_in_cnt = (const int*)B.ccnt[cur];
    ma.carry_out = (uint64_t*)B.carry[cur ^ 1];
    ma.carry_out_cnt = (int*)B.ccnt[cur ^ 1];
    ma.carry_cap = P.carry_cap;
    ma.sub = (const uint64_t*)B.sub;
    ma.sub_cnt = (const int*)B.subcnt;
    ma.sub_cap = P.sub_cap;
    ma.slices = R.slices;
    ma.slices_max = P.slices_max;
    ma.thr = (uint32_t*)B.thr;
    ma.kc = (const int*)B.kc;
    ma.n = (int)n;
    ma.bitonic = L.bitonic;
    // expected union: the first round's rows (nothing is carried into it), else the carry (with
    // the band's rows beyond K') plus about cmax * (new rows / rows seen) candidates
    {
      const long seen = R.begin - P.rounds[0].begin;
      if (seen <= 0) {
        ma.wide = R.end - R.begin > 1024 ? 1 : 0;   // exact: at most one entry per row
      } else {
        const double est = (double)L.cmax * (double)(R.end - P.rounds[0].begin) / (double)seen +
                           (L.band ? 0.5 * L.cmax : 0.0);
        ma.wide = est > 900.0 ? 1 : 0;   // an estimate: keep a margin below the 1024 slots
      }
    }
    ma.flags = (int*)B.flags;
    ma.self_ids = (const int*)B.self;
    ma.n_dev = L.n_dev;
    ma.band = band;
    ma.seen = r + 1 == P.rounds.size()
                  ? 1.0
                  : (double)(R.end - P.rounds[0].begin) / (double)(P.rounds.back().end - P.rounds[0].begin);
    KNNS_TRY(launch_merge(ma, st));
    if (hook) (*hook)(2);
    cur ^= 1;
  }

  FinalArgs fa{};
  fa.carry = (const uint64_t*)B.carry[cur];
  fa.carry_cnt = (const int*)B.ccnt[cur];
  fa.carry_cap = P.carry_cap;
  fa.kc = (const int*)B.kc;
  fa.kq = (const int*)B.kq;
  fa.out_off = (const long*)B.off;
  if (c->dtype == KNNS_F64) {
    fa.Xm = c->x64; fa.ldxm = c->d; fa.Qm = B.q64; fa.ldqm = c->d; fa.master_f64 = 1;
  } else {
    fa.Xm = c->x32; fa.ldxm = c->ld32; fa.Qm = B.q32; fa.ldqm = c->ld32; fa.master_f64 = 0;
  }
  fa.d = c->d;
  fa.out_f64 = L.out64 ? 1 : 0;
  fa.out_dist = B.odist;
  fa.out_idx = (int32_t*)B.oidx;
  fa.n = (int)n;
  fa.n_dev = L.n_dev;
  fa.certify = L.certify && B.qn2 ? 1 : 0;
  fa.rows = std::max<long>(0, (P.rounds.empty() ? 0 : P.rounds.back().end - P.rounds[0].begin) - (L.excl ? 1 : 0));
  fa.thr = (uint32_t*)B.thr;
  fa.flags = (int*)B.flags;
  fa.qn2 = (const double*)B.qn2;
  fa.qrho = (const double*)B.qrho;
  fa.bnd = c->bnd;
  fa.bslot = P.prec_kernel;
  fa.gamma = key_gamma(P);
  fa.prune = band.on;
  fa.rrho = c->rrho && pk != PREC_FP8 ? c->rrho + (pk == PREC_BF16 ? c->rows_pad : 0) : nullptr;
  if (L.feed) KNNS_TRY(feed_rows(*L.feed, c->rows_pad, st));
  if (hook) (*hook)(3);
  KNNS_TRY(launch_final(fa, st));
  if (hook) (*hook)(4);
#undef KNNS_TRY
  return hipSuccess;
}

namespace {

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

}  // namespace

// ---- re-search of uncertified queries ----
knns_status run_research(const ResearchIn& R, Workspace& ws, hipStream_t st, int cus, long* researched) {
  const knns_corpus_s* c = R.c;
  const long nf = (long)R.qids.size();
  *researched = nf;
  if (nf == 0) return KNNS_OK;
  const long m = R.row_end - R.row_begin;
  const long m_eff = R.excl ? std::max<long>(m - 1, 1) : m;
  const int cap = (int)std::max<long>(1, std::min<long>(2048, m_eff));
  int kmax = 1;
  for (int k : R.kq) kmax = std::max(kmax, k);
  const size_t qsz = R.dtype == KNNS_F64 ? 8 : 4;
  const size_t row_bytes = (size_t)c->d * qsz;
  // plan: the preset thresholds replace the dense first round; the slice count is lowered until
  // the workspace fits the provider's budget (the answer does not depend on it)
  const Tuning tu = tuning_for(R.prec, c->d, kmax, nf, m);
  int sreq = 0;
  Plan P;
  SearchSizes z;
  size_t ctl_bytes = 0, total = 0;
  long n_pad = 0;
  for (;;) {
    // few rounds: the thresholds start where the certificate needs them and rarely tighten
    P = make_plan(m, c->d, nf, kmax, cap, R.prec, cus, 64, sreq, tu.min_waves, false, tu.qb);
    z = search_sizes(c, P, nf, 0, R.excl, R.out64);
    n_pad = P.n_pad;
    ctl_bytes = (size_t)nf * (4 + 4 + 8 + 4) + (size_t)n_pad * (4 + 4) + 64;
    total = 0;
    const size_t parts[] = {z.q32, z.q16, z.q8, z.q64, z.sub, z.subcnt, z.carry, z.carry, z.ccnt, z.ccnt,
                            z.flags, z.qn2, z.qrho, (size_t)nf * row_bytes, ctl_bytes};
    for (size_t b : parts) total += (b + 255) & ~(size_t)255;
    if (total <= ws.budget()) break;
    if (P.slices_max <= 1)
      return fail(KNNS_ERR_NOMEM, "re-search of %ld uncertified queries needs %zu bytes of workspace (budget %zu)",
                  nf, total, ws.budget());
    sreq = std::max(1, P.slices_max / 2);
  }
  for (Round& r : P.rounds) {
    r.begin += R.row_begin;
    r.end += R.row_begin;
  }
  // control block: k, shortlist, output offset, source row of each query; threshold and self row of
  // each query slot
  std::vector<char> ctl(ctl_bytes, 0);
  size_t o = 0;
  auto place = [&](size_t bytes) { const size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
  const size_t o_kq = place(nf * 4), o_kc = place(nf * 4), o_off = place(nf * 8), o_idx = place(nf * 4),
               o_thr = place(n_pad * 4), o_self = place(n_pad * 4);
  if (o > ctl_bytes) return fail(KNNS_ERR_INVALID, "internal: control block layout");
  for (long i = 0; i < nf; ++i) {
    ((int*)(ctl.data() + o_kq))[i] = R.kq[i];
    ((int*)(ctl.data() + o_kc))[i] = cap;
    ((long*)(ctl.data() + o_off))[i] = R.off[i];
    ((int*)(ctl.data() + o_idx))[i] = R.qids[i];
    ((uint32_t*)(ctl.data() + o_thr))[i] = R.thr[i];
  }
  for (long i = 0; i < n_pad; ++i) ((int*)(ctl.data() + o_self))[i] = (R.excl && i < nf) ? R.self[i] : -1;
  SearchBuffers B;
  void* rq = nullptr;
  void* dctl = nullptr;
  knns_status sst;
#define KNNS_RWS(slot, bytes, ptr) \
  if ((bytes) && (sst = ws.get(slot, (size_t)(bytes), &ptr)) != KNNS_OK) return sst
  KNNS_RWS(W_Q32, z.q32, B.q32);
  KNNS_RWS(W_Q16, z.q16, B.q16);
  KNNS_RWS(W_Q8, z.q8, B.q8);
  KNNS_RWS(W_Q64, z.q64, B.q64);
  KNNS_RWS(W_SUB, z.sub, B.sub);
  KNNS_RWS(W_SUBCNT, z.subcnt, B.subcnt);
  KNNS_RWS(W_CARRY0, z.carry, B.carry[0]);
  KNNS_RWS(W_CARRY1, z.carry, B.carry[1]);
  KNNS_RWS(W_CCNT0, z.ccnt, B.ccnt[0]);
  KNNS_RWS(W_CCNT1, z.ccnt, B.ccnt[1]);
  KNNS_RWS(W_FLAGS, z.flags, B.flags);
  KNNS_RWS(W_QN2, z.qn2, B.qn2);
  KNNS_RWS(W_QRHO, z.qrho, B.qrho);
  KNNS_RWS(W_RQ, (size_t)nf * row_bytes, rq);
  KNNS_RWS(W_RCTL, ctl_bytes, dctl);
#undef KNNS_RWS
  B.kq = (char*)dctl + o_kq;
  B.kc = (char*)dctl + o_kc;
  B.off = (char*)dctl + o_off;
  B.thr = (char*)dctl + o_thr;
  B.self = R.excl ? (char*)dctl + o_self : nullptr;
  B.odist = R.odist;
  B.oidx = R.oidx;
  if ((sst = ws.upload(dctl, ctl.data(), ctl_bytes, st)) != KNNS_OK) return sst;
  KNNS_HIP(launch_gather_rows(rq, R.qsrc, (const int*)((char*)dctl + o_idx), nf, (long)row_bytes, st));
  SearchLaunch L;
  L.c = c;
  L.P = &P;
  L.n = nf;
  L.dtype = R.dtype;
  L.qsrc = rq;
  L.uniform = false;
  L.kuni = 1;
  L.cmax = cap;
  L.excl = R.excl;
  L.self_preset = true;
  L.thr_preset = true;
  L.certify = true;
  L.out64 = R.out64;
  L.bitonic = R.bitonic ? 1 : 0;
  KNNS_HIP(enqueue_search(L, B, st, nullptr));
  int flags_h[16] = {0};
  KNNS_HIP(hipMemcpyAsync(flags_h, B.flags, sizeof(flags_h), hipMemcpyDeviceToHost, st));
  KNNS_HIP(hipStreamSynchronize(st));
  if (flags_h[0] != 0)
    return fail(KNNS_ERR_OVERFLOW, "candidate pool overflow in the re-search (%d events): too many exact ties", flags_h[0]);
  if (flags_h[2] != 0)
    return fail(KNNS_ERR_OVERFLOW, "%d queries cannot be certified: more than %d corpus rows lie within the "
                "distance error bound of their k-th neighbour", flags_h[2], cap);
  return KNNS_OK;
}

namespace {

// The engine's own workspace for a re-search: its grow-only slots and pinned staging.
struct EngineWorkspace : Workspace {
  Engine& E;
  explicit EngineWorkspace(Engine& e) : E(e) {}
  size_t budget() const override { return (size_t)4 << 30; }
  knns_status get(int slot, size_t bytes, void** out) override { return E.get((Slot)slot, bytes, out); }
  knns_status upload(void* dst, const void* src, size_t bytes, hipStream_t st) override {
    void* pin = nullptr;
    if (knns_status s = E.pinned(bytes, &pin)) return s;
    std::memcpy(pin, src, bytes);
    KNNS_HIP(hipMemcpyAsync(dst, pin, bytes, hipMemcpyHostToDevice, st));
    return KNNS_OK;
  }
};

}  // namespace

knns_status research_on_engine(const ResearchIn& R, hipStream_t st, long* researched) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  if (knns_status s = E.init()) return s;
  EngineWorkspace ws(E);
  return run_research(R, ws, st, E.cus, researched);
}

namespace {

// Searches corpus rows [row_begin, row_end) (row_end < 0: to the last row); indices are corpus rows.
// self_host (optional, n entries): corpus row query i never returns (< 0: none); it replaces
// KNNS_EXCLUDE_SELF's self_offset + i.
knns_status do_search(knns_corpus c, const void* queries, int dtype, long n, const int* kq_host,
                      int kuni, const knns_params* p, void* dist, int32_t* idx,
                      long row_begin = 0, long row_end = -1, Upload* feed = nullptr,
                      const long* self_host = nullptr) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  g_err.clear();
  if (knns_status st = E.init()) return st;
  if (!c) return fail(KNNS_ERR_INVALID, "null corpus");
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  if (n < 0) return fail(KNNS_ERR_INVALID, "negative query count %ld", n);
  if (n == 0) return KNNS_OK;
  if (!queries || !dist || !idx) return fail(KNNS_ERR_INVALID, "null query/output pointer");
  if (dtype != c->dtype)
    return fail(KNNS_ERR_INVALID, "query dtype %d does not match corpus dtype %d", dtype, c->dtype);
  if (n > (1L << 30)) return fail(KNNS_ERR_INVALID, "too many queries (%ld)", n);
  const int prec = p->precision;
  if (prec != KNNS_PREC_FP32 && prec != KNNS_PREC_FP8 && prec != KNNS_PREC_BF16 && prec != KNNS_PREC_FP64)
    return fail(KNNS_ERR_INVALID, "unknown precision %d", prec);
  if (prec == KNNS_PREC_FP64 && c->dtype != KNNS_F64)
    return fail(KNNS_ERR_INVALID, "FP64 precision needs a double corpus");
  if (prec == KNNS_PREC_BF16 && !c->x16)
    return fail(KNNS_ERR_INVALID, "BF16 precision needs a corpus built with KNNS_CORPUS_BF16");
  if (prec == KNNS_PREC_FP8 && !c->x8)
    return fail(KNNS_ERR_INVALID, "FP8 precision needs a corpus built with KNNS_CORPUS_FP8");
  const bool dev = (p->flags & KNNS_DEVICE_PTRS) != 0;
  const bool excl = (p->flags & KNNS_EXCLUDE_SELF) != 0 || self_host != nullptr;
  const bool out64 = c->dtype == KNNS_F64 || (p->flags & KNNS_DIST_F64);
  const bool ktimes = (p->flags & KNNS_KERNEL_TIMES) != 0;
  hipStream_t st = p->stream ? (hipStream_t)p->stream : E.stream;
  if (row_end < 0) row_end = c->m;
  if (row_begin < 0 || row_begin >= row_end || row_end > c->m)
    return fail(KNNS_ERR_INVALID, "row range [%ld, %ld) outside the corpus' %ld rows", row_begin, row_end,
                c->m);
  const long m = row_end - row_begin;   // rows searched

  // per-query k and shortlist
  const bool uniform = kq_host == nullptr;
  const long m_eff = excl ? std::max<long>(m - 1, 1) : m;
  // the tuning entry of the search (its largest k), then the shortlists
  int kmax_all = uniform ? kuni : 1;
  if (!uniform)
    for (long i = 0; i < n; ++i) kmax_all = std::max(kmax_all, kq_host[i]);
  const Tuning tu = tuning_for(prec, c->d, std::min(std::max(kmax_all, 1), 1024), n, m);
  auto shortlist_of = [&](int k) { return shortlist_for(prec, k, p->shortlist, tu.sl_extra, m_eff); };
  std::vector<int> kq, kc;
  std::vector<long> off;
  long total_out = 0;
  int kmax = 0, cmax = 0;
  if (uniform) {
    if (kuni < 1 || kuni > 1024) return fail(KNNS_ERR_INVALID, "k=%d outside [1, 1024]", kuni);
    kmax = kuni;
    cmax = shortlist_of(kuni);
    total_out = n * (long)kuni;
  } else {
    kq.resize(n);
    kc.resize(n);
    off.resize(n);
    for (long i = 0; i < n; ++i) {
      const int k = kq_host[i];
      if (k < 1 || k > 1024) return fail(KNNS_ERR_INVALID, "k=%d outside [1, 1024]", k);
      kq[i] = k;
      kc[i] = shortlist_of(k);
      off[i] = total_out;
      total_out += k;
      kmax = std::max(kmax, k);
      cmax = std::max(cmax, kc[i]);
    }
  }

  const bool band = (p->flags & KNNS_NO_CERTIFY) == 0 && prec != KNNS_PREC_FP8;
  Plan P = make_plan(m, c->d, n, kmax, cmax, prec, E.cus, p->ratio > 0 ? p->ratio : tu.ratio, p->slices,
                     tu.min_waves, true, p->qb > 0 ? p->qb : tu.qb, band);
  for (Round& r : P.rounds) {   // planned over the range's rows, from its first row
    r.begin += row_begin;
    r.end += row_begin;
  }
  if (P.dp != c->dp) return fail(KNNS_ERR_INVALID, "corpus/plan dimension mismatch");

  // workspace (grow-only engine slots)
  const SearchSizes z = search_sizes(c, P, n, total_out, excl, out64);
  SearchBuffers B;
  knns_status sst;
#define KNNS_WS(slot, bytes, ptr) \
  if ((sst = E.get(slot, (size_t)(bytes), &ptr)) != KNNS_OK) return sst
  KNNS_WS(W_Q32, z.q32, B.q32);
  if (z.q16) KNNS_WS(W_Q16, z.q16, B.q16);
  if (z.q8) KNNS_WS(W_Q8, z.q8, B.q8);
  if (z.q64) KNNS_WS(W_Q64, z.q64, B.q64);
  KNNS_WS(W_THR, z.thr, B.thr);
  KNNS_WS(W_KC, z.kc, B.kc);
  KNNS_WS(W_KQ, z.kq, B.kq);
  KNNS_WS(W_OFF, z.off, B.off);
  if (z.self) KNNS_WS(W_SELF, z.self, B.self);
  KNNS_WS(W_SUB, z.sub, B.sub);
  KNNS_WS(W_SUBCNT, z.subcnt, B.subcnt);
  KNNS_WS(W_CARRY0, z.carry, B.carry[0]);
  KNNS_WS(W_CARRY1, z.carry, B.carry[1]);
  KNNS_WS(W_CCNT0, z.ccnt, B.ccnt[0]);
  KNNS_WS(W_CCNT1, z.ccnt, B.ccnt[1]);
  KNNS_WS(W_FLAGS, z.flags, B.flags);
  KNNS_WS(W_QN2, z.qn2, B.qn2);
  KNNS_WS(W_QRHO, z.qrho, B.qrho);
  if (!dev) {
    KNNS_WS(W_ODIST, z.odist, B.odist);
    KNNS_WS(W_OIDX, z.oidx, B.oidx);
  } else {
    B.odist = dist;
    B.oidx = idx;
  }

  size_t ev = 0;
  hipEvent_t e_begin = E.event(ev++);
  if (!e_begin) return fail(KNNS_ERR_DEVICE, "event creation failed");
  KNNS_HIP(hipEventRecord(e_begin, st));

  const size_t qsz = dtype == KNNS_F64 ? 8 : 4;
  const void* qsrc = queries;
  if (!dev) {
    void* stage;
    KNNS_WS(W_STAGE, n * (long)c->d * qsz, stage);
    KNNS_HIP(hipMemcpyAsync(stage, queries, n * (long)c->d * qsz, hipMemcpyHostToDevice, st));
    qsrc = stage;
  }
#undef KNNS_WS
  // per-query k and self rows: staged through pinned host memory (no synchronisation inside the
  // search). Layout: k and shortlist (4 bytes each per query), padded to 8 bytes, the offsets
  // (8 bytes each), then the self rows of every query slot (4 bytes each, -1 past n).
  if (!uniform || self_host) {
    void* pin;
    if ((sst = E.pinned(n * 16 + 8 + (long)P.n_pad * 4, &pin)) != KNNS_OK) return sst;
    int* pk = (int*)pin;
    int* pc = pk + n;
    long* po = (long*)(pc + n + (n & 1));
    int* ps = (int*)(po + n);
    if (!uniform) {
      std::memcpy(pk, kq.data(), n * 4);
      std::memcpy(pc, kc.data(), n * 4);
      std::memcpy(po, off.data(), n * 8);
      KNNS_HIP(hipMemcpyAsync(B.kq, pk, n * 4, hipMemcpyHostToDevice, st));
      KNNS_HIP(hipMemcpyAsync(B.kc, pc, n * 4, hipMemcpyHostToDevice, st));
      KNNS_HIP(hipMemcpyAsync(B.off, po, n * 8, hipMemcpyHostToDevice, st));
    }
    if (self_host) {
      // self rows are searched-range rows: a row outside [row_begin, row_end) is never returned
      for (long i = 0; i < P.n_pad; ++i) {
        const long r = i < n ? self_host[i] : -1;
        ps[i] = (r >= row_begin && r < row_end) ? (int)r : -1;
      }
      KNNS_HIP(hipMemcpyAsync(B.self, ps, (size_t)P.n_pad * 4, hipMemcpyHostToDevice, st));
    }
  }

  SearchLaunch L;
  L.c = c;
  L.P = &P;
  L.n = n;
  L.dtype = dtype;
  L.qsrc = qsrc;
  L.uniform = uniform;
  L.kuni = uniform ? kuni : 1;
  L.cmax = cmax;
  L.excl = excl;
  L.self_offset = p->self_offset;
  L.self_preset = self_host != nullptr;
  L.out64 = out64;
  L.bitonic = p->select == KNNS_SELECT_BITONIC ? 1 : 0;
  L.feed = feed;
  L.certify = (p->flags & KNNS_NO_CERTIFY) == 0;
  L.band = band;
  // per-launch events (KNNS_KERNEL_TIMES): pairs (scan begin, scan end), (scan end, merge end),
  // (final begin, final end)
  std::vector<std::pair<size_t, size_t>> scan_ev, sel_ev;
  size_t last_scan_end = 0, final_begin = 0;
  bool ev_ok = true;
  LaunchHook hook = [&](int what) {
    const size_t i = ev++;
    hipEvent_t e = E.event(i);
    if (!e || hipEventRecord(e, st) != hipSuccess) { ev_ok = false; return; }
    if (what == 0) scan_ev.push_back({i, 0});
    else if (what == 1) { scan_ev.back().second = i; last_scan_end = i; }
    else if (what == 2) sel_ev.push_back({last_scan_end, i});
    else if (what == 3) final_begin = i;
    else sel_ev.push_back({final_begin, i});
  };
  KNNS_HIP(enqueue_search(L, B, st, ktimes ? &hook : nullptr));
  if (!ev_ok) return fail(KNNS_ERR_DEVICE, "kernel timing event failed");

  int flags_h[16] = {0};
  KNNS_HIP(hipMemcpyAsync(flags_h, B.flags, sizeof(flags_h), hipMemcpyDeviceToHost, st));
  if (!dev) {
    KNNS_HIP(hipMemcpyAsync(dist, B.odist, z.odist, hipMemcpyDeviceToHost, st));
    KNNS_HIP(hipMemcpyAsync(idx, B.oidx, z.oidx, hipMemcpyDeviceToHost, st));
  }
  size_t e_end = ev++;
  KNNS_HIP(hipEventRecord(E.event(e_end), st));
  KNNS_HIP(hipStreamSynchronize(st));

  // Queries whose shortlist the final kernel could not certify are searched again with every row
  // below their certificate threshold admitted (exact forms; FP8 is the approximate form and only
  // reports them).
  long researched = 0;
  double research_s = 0.0;
  if (flags_h[2] > 0 && flags_h[0] == 0 && L.certify && prec != KNNS_PREC_FP8) {
    ResearchIn R;
    R.c = c;
    R.prec = prec;
    R.dtype = dtype;
    R.out64 = out64;
    R.excl = excl;
    R.bitonic = L.bitonic != 0;
    R.row_begin = row_begin;
    R.row_end = row_end;
    R.qsrc = qsrc;
    R.odist = B.odist;
    R.oidx = (int32_t*)B.oidx;
    std::vector<uint32_t> need(n);
    KNNS_HIP(hipMemcpy(need.data(), B.thr, n * 4, hipMemcpyDeviceToHost));
    for (long i = 0; i < n; ++i) {
      if (!need[i]) continue;
      R.qids.push_back((int)i);
      R.thr.push_back(need[i]);
      R.kq.push_back(uniform ? kuni : kq[i]);
      R.off.push_back(uniform ? i * (long)kuni : off[i]);
      long sr = -1;
      if (self_host) sr = self_host[i];
      else if (excl) sr = p->self_offset + i;
      R.self.push_back((sr >= row_begin && sr < row_end && sr < c->m) ? (int)sr : -1);
    }
    EngineWorkspace ws(E);
    const size_t r_begin = ev++;
    if (ktimes) KNNS_HIP(hipEventRecord(E.event(r_begin), st));
    knns_status rs = run_research(R, ws, st, E.cus, &researched);
    if (rs == KNNS_ERR_OVERFLOW && prec == KNNS_PREC_BF16) {
      // more than the re-search's shortlist of rows lie within the BF16 error bound of some of
      // these queries' k-th neighbour: search them once more on the FP32 operand rows (an error
      // bound 2^16 times smaller) with every row admitted
      g_err.clear();
      R.prec = KNNS_PREC_FP32;
      std::fill(R.thr.begin(), R.thr.end(), ORD_POS_INF);
      rs = run_research(R, ws, st, E.cus, &researched);
    }
    if (rs) return rs;
    if (ktimes) {
      const size_t r_end = ev++;
      KNNS_HIP(hipEventRecord(E.event(r_end), st));
      KNNS_HIP(hipEventSynchronize(E.event(r_end)));
      float rms = 0.f;
      KNNS_HIP(hipEventElapsedTime(&rms, E.event(r_begin), E.event(r_end)));
      research_s = rms * 1e-3;
    }
    if (!dev) {
      KNNS_HIP(hipMemcpyAsync(dist, B.odist, z.odist, hipMemcpyDeviceToHost, st));
      KNNS_HIP(hipMemcpyAsync(idx, B.oidx, z.oidx, hipMemcpyDeviceToHost, st));
    }
    e_end = ev++;
    KNNS_HIP(hipEventRecord(E.event(e_end), st));
    KNNS_HIP(hipStreamSynchronize(st));
  }

  knns_stats S{};
  float ms = 0.f;
  KNNS_HIP(hipEventElapsedTime(&ms, e_begin, E.event(e_end)));
  S.total_s = ms * 1e-3;
  for (auto& pr : scan_ev) {
    KNNS_HIP(hipEventElapsedTime(&ms, E.event(pr.first), E.event(pr.second)));
    S.scan_s += ms * 1e-3;
  }
  for (auto& pr : sel_ev) {
    KNNS_HIP(hipEventElapsedTime(&ms, E.event(pr.first), E.event(pr.second)));
    S.select_s += ms * 1e-3;
  }
  S.rounds = (long)P.rounds.size();
  S.slow_compactions = flags_h[1];
  S.groups = P.groups;
  S.slices_max = P.slices_max;
  S.qb = P.qb;
  S.dc = P.dc;
  S.dp = P.dp;
  S.uncertified = flags_h[2];
  S.shortlist = cmax;
  S.researched = researched;
  S.research_s = research_s;
  E.last = S;
  if (flags_h[0] != 0)
    return fail(KNNS_ERR_OVERFLOW, "candidate pool overflow (%d events): too many exact ties",
                flags_h[0]);
  return KNNS_OK;
}

// ---- search without allocation (knns_search_noalloc) ----
// Every workspace buffer of the search is carved from `scratch`, a file-scope device array of
// KNNS_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's load until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress, with that device
// current) and kept in g_scratch_ptr: host code never passes the array's name to a kernel or a
// copy, only that address.
static __device__ __align__(256) unsigned char scratch[KNNS_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_cus[kMaxDevices] = {};

knns_status scratch_for_device(int dev, void** ptr, int* cus) {
  if (dev < 0 || dev >= kMaxDevices) return fail(KNNS_ERR_UNSUPPORTED, "device %d outside the first %d", dev, kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void* sp = nullptr;
    const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
      (void)hipGetLastError();
      return fail(KNNS_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
      (void)hipGetLastError();
      return fail(KNNS_ERR_DEVICE, "device query failed");
    }
    g_scratch_cus[dev] = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    g_scratch_ptr[dev] = sp;
  }
  *ptr = g_scratch_ptr[dev];
  *cus = g_scratch_cus[dev];
  return KNNS_OK;
}

// The scratch array as a re-search workspace: buffers carved in order, copies from pageable host
// memory (no pinned staging).
struct ScratchWorkspace : Workspace {
  char* base;
  size_t cap, used = 0;
  ScratchWorkspace(void* b, size_t c) : base((char*)b), cap(c) {}
  size_t budget() const override { return cap; }
  knns_status get(int, size_t bytes, void** out) override {
    const size_t b = (bytes + 255) & ~(size_t)255;
    if (used + b > cap) return fail(KNNS_ERR_NOMEM, "re-search workspace exceeds the scratch array");
    *out = base + used;
    used += b;
    return KNNS_OK;
  }
  knns_status upload(void* dst, const void* src, size_t bytes, hipStream_t st) override {
    KNNS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return KNNS_OK;
  }
};

// Bytes of the carved workspace (each buffer aligned to 256 bytes).
size_t carve_bytes(const SearchSizes& z) {
  const size_t parts[] = {z.q32, z.q16, z.q8, z.q64, z.thr, z.kc, z.kq, z.off, z.self, z.sub, z.subcnt,
                          z.carry, z.carry, z.ccnt, z.ccnt, z.flags, z.qn2, z.qrho};
  size_t t = 0;
  for (size_t b : parts) t += (b + 255) & ~(size_t)255;
  return t;
}

knns_status do_search_noalloc(knns_corpus c, const void* queries, int dtype, long n, const knns_params* p,
                              void* dist, int32_t* idx) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);   // calls of one process never overlap
  g_err.clear();
  if (!c) return fail(KNNS_ERR_INVALID, "null corpus");
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  if (n < 0) return fail(KNNS_ERR_INVALID, "negative query count %ld", n);
  if (n == 0) return KNNS_OK;
  if (!queries || !dist || !idx) return fail(KNNS_ERR_INVALID, "null query/output pointer");
  if (!(p->flags & KNNS_DEVICE_PTRS))
    return fail(KNNS_ERR_INVALID, "knns_search_noalloc takes device pointers (KNNS_DEVICE_PTRS)");
  if (p->flags & KNNS_KERNEL_TIMES) return fail(KNNS_ERR_INVALID, "knns_search_noalloc records no kernel times");
  if (dtype != c->dtype)
    return fail(KNNS_ERR_INVALID, "query dtype %d does not match corpus dtype %d", dtype, c->dtype);
  if (n > (1L << 30)) return fail(KNNS_ERR_INVALID, "too many queries (%ld)", n);
  const int prec = p->precision;
  if (prec != KNNS_PREC_FP32 && prec != KNNS_PREC_FP8 && prec != KNNS_PREC_BF16 && prec != KNNS_PREC_FP64)
    return fail(KNNS_ERR_INVALID, "unknown precision %d", prec);
  if (prec == KNNS_PREC_FP64 && c->dtype != KNNS_F64) return fail(KNNS_ERR_INVALID, "FP64 precision needs a double corpus");
  if (prec == KNNS_PREC_BF16 && !c->x16)
    return fail(KNNS_ERR_INVALID, "BF16 precision needs a corpus built with KNNS_CORPUS_BF16");
  if (prec == KNNS_PREC_FP8 && !c->x8) return fail(KNNS_ERR_INVALID, "FP8 precision needs a corpus built with KNNS_CORPUS_FP8");
  const int k = p->k;
  if (k < 1 || k > 1024) return fail(KNNS_ERR_INVALID, "k=%d outside [1, 1024]", k);
  const bool excl = (p->flags & KNNS_EXCLUDE_SELF) != 0;
  const bool out64 = c->dtype == KNNS_F64 || (p->flags & KNNS_DIST_F64);
  int dev = 0;
  KNNS_HIP(hipGetDevice(&dev));
  void* base = nullptr;
  int cus = 256;
  if (knns_status st = scratch_for_device(dev, &base, &cus)) return st;
  const long m = c->m;
  const long m_eff = excl ? std::max<long>(m - 1, 1) : m;
  const Tuning tu = tuning_for(prec, c->d, k, n, m);
  const int cmax = shortlist_for(prec, k, p->shortlist, tu.sl_extra, m_eff);
  const bool band = (p->flags & KNNS_NO_CERTIFY) == 0 && prec != KNNS_PREC_FP8;
  // the planner's slice count, lowered until the workspace fits the scratch array (the answer does
  // not depend on the slice count, only the parallelism of the scan does)
  int sreq = p->slices;
  Plan P;
  SearchSizes z;
  for (;;) {
    P = make_plan(m, c->d, n, k, cmax, prec, cus, p->ratio > 0 ? p->ratio : tu.ratio, sreq, tu.min_waves, true,
                  p->qb > 0 ? p->qb : tu.qb, band);
    z = search_sizes(c, P, n, n * (long)k, excl, out64);
    if (carve_bytes(z) <= (size_t)KNNS_NOALLOC_SCRATCH_BYTES) break;
    if (P.slices_max <= 1)
      return fail(KNNS_ERR_NOMEM,
                  "search needs %zu bytes of workspace even with one slice (%ld queries, k=%d): above the %ld-byte "
                  "scratch of knns_search_noalloc; use knns_search or fewer queries per call",
                  carve_bytes(z), n, k, (long)KNNS_NOALLOC_SCRATCH_BYTES);
    sreq = std::max(1, P.slices_max / 2);
  }
  if (P.dp != c->dp) return fail(KNNS_ERR_INVALID, "corpus/plan dimension mismatch");
  SearchBuffers B;
  char* cur = (char*)base;
  auto take = [&](size_t bytes) -> void* {
    if (!bytes) return nullptr;
    void* r = cur;
    cur += (bytes + 255) & ~(size_t)255;
    return r;
  };
  B.q32 = take(z.q32);
  B.q16 = take(z.q16);
  B.q8 = take(z.q8);
  B.q64 = take(z.q64);
  B.thr = take(z.thr);
  B.kc = take(z.kc);
  B.kq = take(z.kq);
  B.off = take(z.off);
  B.self = take(z.self);
  B.sub = take(z.sub);
  B.subcnt = take(z.subcnt);
  B.carry[0] = take(z.carry);
  B.carry[1] = take(z.carry);
  B.ccnt[0] = take(z.ccnt);
  B.ccnt[1] = take(z.ccnt);
  B.flags = take(z.flags);
  B.qn2 = take(z.qn2);
  B.qrho = take(z.qrho);
  B.odist = dist;
  B.oidx = idx;
  hipStream_t st = (hipStream_t)p->stream;   // NULL: the device's null stream
  SearchLaunch L;
  L.c = c;
  L.P = &P;
  L.n = n;
  L.dtype = dtype;
  L.qsrc = queries;
  L.uniform = true;
  L.kuni = k;
  L.cmax = cmax;
  L.excl = excl;
  L.self_offset = p->self_offset;
  L.out64 = out64;
  L.bitonic = p->select == KNNS_SELECT_BITONIC ? 1 : 0;
  L.certify = (p->flags & KNNS_NO_CERTIFY) == 0;
  L.band = band;
  // the flags (overflow counter) are cleared by a fill kernel inside the sequence
  KNNS_HIP(enqueue_search(L, B, st, nullptr));
  int flags_h[16] = {0};
  KNNS_HIP(hipMemcpyAsync(flags_h, B.flags, sizeof(flags_h), hipMemcpyDeviceToHost, st));
  KNNS_HIP(hipStreamSynchronize(st));
  if (flags_h[0] != 0)
    return fail(KNNS_ERR_OVERFLOW, "candidate pool overflow (%d events): too many exact ties", flags_h[0]);
  if (flags_h[2] > 0 && L.certify && prec != KNNS_PREC_FP8) {
    // uncertified queries: re-searched inside the same scratch array (the search above is done)
    ResearchIn R;
    R.c = c;
    R.prec = prec;
    R.dtype = dtype;
    R.out64 = out64;
    R.excl = excl;
    R.bitonic = L.bitonic != 0;
    R.row_begin = 0;
    R.row_end = m;
    R.qsrc = queries;
    R.odist = dist;
    R.oidx = idx;
    std::vector<uint32_t> need(n);
    KNNS_HIP(hipMemcpyAsync(need.data(), B.thr, n * 4, hipMemcpyDeviceToHost, st));
    KNNS_HIP(hipStreamSynchronize(st));
    for (long i = 0; i < n; ++i) {
      if (!need[i]) continue;
      R.qids.push_back((int)i);
      R.thr.push_back(need[i]);
      R.kq.push_back(k);
      R.off.push_back(i * (long)k);
      const long sr = excl ? p->self_offset + i : -1;
      R.self.push_back((sr >= 0 && sr < m) ? (int)sr : -1);
    }
    ScratchWorkspace ws(base, (size_t)KNNS_NOALLOC_SCRATCH_BYTES);
    long researched = 0;
    return run_research(R, ws, st, cus, &researched);
  }
  return KNNS_OK;
}

}  // namespace
}  // namespace knns

using namespace knns;

extern "C" {

knns_status knns_search_noalloc(knns_corpus c, const void* queries, int dtype, long n,
                                const knns_params* p, void* dist, int32_t* idx) {
  return do_search_noalloc(c, queries, dtype, n, p, dist, idx);
}

void knns_params_init(knns_params* p, int precision, int k) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->precision = precision;
  p->select = KNNS_SELECT_AUTO;
  p->k = k;
}

}  // extern "C"

namespace knns {
namespace {

// Validation and device buffers of a corpus; its rows are written by an Upload.
knns_status corpus_alloc(knns_corpus* out, const void* data, int dtype, long m, int d,
                         unsigned flags) {
  Engine& E = engine();
  if (!out) return fail(KNNS_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (!data) return fail(KNNS_ERR_INVALID, "null corpus data");
  if (m < 1 || m > 0x7fffff00L) return fail(KNNS_ERR_INVALID, "corpus rows %ld out of range", m);
  if (d < 1 || d > 65536) return fail(KNNS_ERR_INVALID, "dimension %d out of range", d);
  if (dtype != KNNS_F32 && dtype != KNNS_F64) return fail(KNNS_ERR_INVALID, "bad dtype %d", dtype);
  if ((flags & KNNS_CORPUS_FP8) && fp8_row_bytes(d) > 256)
    return fail(KNNS_ERR_UNSUPPORTED, "FP8 corpus copy supports d <= 256 (d=%d)", d);
  if (knns_status st = E.init()) return st;
  knns_corpus c = new knns_corpus_s();
  c->m = m;
  c->d = d;
  c->dtype = dtype;
  c->flags = flags;
  dims_for(d, &c->dp, &c->dc);
  // whole scan stages (64 or 128 rows) never leave the padding, also for a search over a row
  // range (knns_search_rows) whose slices start at any row: 128 rows of slack past round_up(m, 128)
  c->rows_pad = round_up(m, 128) + 128;
  // FP32 rows [dp dims | hi, lo, 0, 0] on a 64-byte stride: the exact rerank gathers a row's first
  // d values from whole 64-byte sectors (8 per 128-dim row instead of 9 on a 528-byte stride)
  c->ld32 = round_up(c->dp + 4, 16);
  c->ld16 = c->dp + 16;
  if (flags & KNNS_CORPUS_FP8) {
    c->dp8 = fp8_row_bytes(d);
    c->ld8 = c->dp8 + 16;
  }
  auto cleanup = [&](knns_status st) {
    if (c->bnd) (void)hipFree(c->bnd);
    if (c->rrho) (void)hipFree(c->rrho);
    if (c->x32) (void)hipFree(c->x32);
    if (c->x16) (void)hipFree(c->x16);
    if (c->x8) (void)hipFree(c->x8);
    if (c->x64) (void)hipFree(c->x64);
    delete c;
    (void)hipGetLastError();
    return st;
  };
  if (hipMalloc((void**)&c->x32, (size_t)c->rows_pad * c->ld32 * 4) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus FP32 rows (%ld x %ld)", c->rows_pad, c->ld32));
  if ((flags & KNNS_CORPUS_BF16) &&
      hipMalloc((void**)&c->x16, (size_t)c->rows_pad * c->ld16 * 2) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus BF16 rows"));
  if ((flags & KNNS_CORPUS_FP8) && hipMalloc((void**)&c->x8, (size_t)c->rows_pad * c->ld8) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus FP8 rows"));
  if (dtype == KNNS_F64 && hipMalloc((void**)&c->x64, (size_t)m * d * 8) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus FP64 master"));
  // certificate maxima, cleared on the upload stream and raised by the pack kernels
  if (hipMalloc((void**)&c->bnd, BND_SLOTS * sizeof(uint32_t)) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus bounds"));
  if (hipMalloc((void**)&c->rrho, 2 * (size_t)c->rows_pad * sizeof(float)) != hipSuccess)
    return cleanup(fail(KNNS_ERR_NOMEM, "corpus row radii"));
  *out = c;
  return KNNS_OK;
}

}  // namespace
}  // namespace knns

extern "C" {

knns_status knns_corpus_create(knns_corpus* out, const void* data, int dtype, long m, int d,
                               unsigned flags) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  g_err.clear();
  knns_corpus c = nullptr;
  if (knns_status st = corpus_alloc(&c, data, dtype, m, d, flags)) return st;
  Upload u;
  knns_status st = u.start(c, data, (flags & KNNS_CORPUS_DEVICE_PTR) != 0, E.stream);
  while (st == KNNS_OK && u.rows_done() < c->rows_pad) st = u.next();
  const knns_status fs = u.finish();
  if (st == KNNS_OK) st = fs;
  if (st != KNNS_OK) {
    knns_corpus_destroy(c);
    return st;
  }
  *out = c;
  return KNNS_OK;
}

knns_status knns_corpus_destroy(knns_corpus c) {
  if (!c) return KNNS_OK;
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  if (c->bnd) (void)hipFree(c->bnd);
  if (c->rrho) (void)hipFree(c->rrho);
  if (c->x32) (void)hipFree(c->x32);
  if (c->x16) (void)hipFree(c->x16);
  if (c->x8) (void)hipFree(c->x8);
  if (c->x64) (void)hipFree(c->x64);
  delete c;
  return KNNS_OK;
}

long knns_corpus_rows(knns_corpus c) { return c ? c->m : 0; }
int knns_corpus_dim(knns_corpus c) { return c ? c->d : 0; }

knns_status knns_search(knns_corpus c, const void* queries, int dtype, long n,
                        const knns_params* p, void* dist, int32_t* idx) {
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  return do_search(c, queries, dtype, n, nullptr, p->k, p, dist, idx);
}

knns_status knns_search_multi(knns_corpus c, const void* queries, int dtype, long n,
                              const int* kq, const knns_params* p, void* dist, int32_t* idx) {
  if (!kq && n > 0) return fail(KNNS_ERR_INVALID, "null per-query k array");
  return do_search(c, queries, dtype, n, kq, 0, p, dist, idx);
}

knns_status knns_search_batches(knns_corpus c, const void* queries, int dtype, long nb,
                                const long* sizes, const int* ks, const long* self_rows,
                                const knns_params* p, void* dist, int32_t* idx) {
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  if (nb < 0) return fail(KNNS_ERR_INVALID, "negative batch count %ld", nb);
  if (nb > 0 && !sizes) return fail(KNNS_ERR_INVALID, "null batch size array");
  long n = 0;
  for (long b = 0; b < nb; ++b) {
    if (sizes[b] < 0) return fail(KNNS_ERR_INVALID, "batch %ld has %ld queries", b, sizes[b]);
    n += sizes[b];
    if (n > (1L << 30)) return fail(KNNS_ERR_INVALID, "too many queries");
  }
  std::vector<int> kq(n);
  std::vector<long> self;
  if (self_rows) self.resize(n);
  long i = 0;
  for (long b = 0; b < nb; ++b)
    for (long j = 0; j < sizes[b]; ++j, ++i) {
      kq[i] = ks ? ks[b] : p->k;
      if (self_rows) self[i] = self_rows[b] < 0 ? -1 : self_rows[b] + j;
    }
  if (n == 0) return KNNS_OK;
  return do_search(c, queries, dtype, n, kq.data(), 0, p, dist, idx, 0, -1, nullptr,
                   self_rows ? self.data() : nullptr);
}

knns_status knns_search_rows(knns_corpus c, long row_begin, long row_end, const void* queries,
                             int dtype, long n, const knns_params* p, void* dist, int32_t* idx) {
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  if (!c) return fail(KNNS_ERR_INVALID, "null corpus");
  if (row_begin < 0 || row_begin >= row_end || row_end > c->m)
    return fail(KNNS_ERR_INVALID, "row range [%ld, %ld) outside the corpus' %ld rows", row_begin, row_end,
                c->m);
  return do_search(c, queries, dtype, n, nullptr, p->k, p, dist, idx, row_begin, row_end);
}

knns_status knns_search_once(const void* data, int dtype, long m, int d, const void* queries,
                             long n, const knns_params* p, void* dist, int32_t* idx) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  g_err.clear();
  if (!p) return fail(KNNS_ERR_INVALID, "null params");
  if (p->flags & KNNS_DEVICE_PTRS) return fail(KNNS_ERR_INVALID, "knns_search_once takes host buffers");
  knns_corpus c = nullptr;
  // the low-precision operand copy the search reads, packed with the FP32 rows as they arrive
  const unsigned cflags = p->precision == KNNS_PREC_BF16 ? KNNS_CORPUS_BF16
                          : (p->precision == KNNS_PREC_FP8 ? KNNS_CORPUS_FP8 : 0u);
  if (knns_status st = corpus_alloc(&c, data, dtype, m, d, cflags)) return st;
  // the corpus goes up on the upload stream; each scan round waits only for the rows it reads
  Upload u;
  if (!E.ustream && hipStreamCreateWithFlags(&E.ustream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    E.ustream = nullptr;
    knns_corpus_destroy(c);
    return fail(KNNS_ERR_DEVICE, "upload stream creation failed");
  }
  knns_status st = u.start(c, data, false, E.ustream);
  if (st == KNNS_OK) st = u.next();
  if (st == KNNS_OK) st = do_search(c, queries, dtype, n, nullptr, p->k, p, dist, idx, 0, -1, &u);
  const knns_status fs = u.finish();
  if (st == KNNS_OK) st = fs;
  knns_corpus_destroy(c);
  return st;
}

knns_status knns_plan_describe(long m, int d, long n, int k, int precision, knns_stats* out) {
  if (!out || m < 1 || d < 1 || n < 1 || k < 1 || k > 1024)
    return fail(KNNS_ERR_INVALID, "bad plan arguments");
  const Tuning tu = tuning_for(precision, d, k, n, m);
  const int cl = shortlist_for(precision, k, 0, tu.sl_extra, m);
  Plan P = make_plan(m, d, n, k, cl, precision, 256, tu.ratio, 0, tu.min_waves, true, tu.qb,
                     precision != KNNS_PREC_FP8);   // as knns_search plans it (certified by default)
  std::memset(out, 0, sizeof(*out));
  out->rounds = (long)P.rounds.size();
  out->groups = P.groups;
  out->slices_max = P.slices_max;
  out->qb = P.qb;
  out->dc = P.dc;
  out->dp = P.dp;
  out->shortlist = cl;
  return KNNS_OK;
}

int knns_predicted_rank(int shortlist, double seen) { return shortlist < 1 ? 0 : predicted_rank(shortlist, seen); }

knns_status knns_set_tuning_table(const char* path) {
  g_err.clear();
  std::vector<TuneEntry> v;
  if (path && path[0]) {
    if (knns_status st = load_tuning(path, &v)) return st;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune = std::move(v);
  g_tune_loaded = true;
  return KNNS_OK;
}

const char* knns_last_error(void) { return g_err.c_str(); }

knns_status knns_last_stats(knns_stats* out) {
  if (!out) return KNNS_ERR_INVALID;
  *out = engine().last;
  return KNNS_OK;
}

int knns_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

void knns_release_workspace(void) {
  Engine& E = engine();
  std::lock_guard<std::recursive_mutex> lock(E.mu);
  E.release();
}

}  // extern "C"

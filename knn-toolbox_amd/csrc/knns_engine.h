// Host-side internals shared by the engine (knns_engine.hip) and the serving loop
// (knns_server.hip). Not part of the ABI.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "../../include/knns_native.h"
#include "knns_internal.h"

// A resident corpus: packed operand rows (and the FP64 master for double corpora).
struct knns_corpus_s {
  long m = 0;
  int d = 0;
  int dp = 0, dc = 0;
  long rows_pad = 0;
  long ld32 = 0, ld16 = 0;
  int dtype = KNNS_F32;
  unsigned flags = 0;
  float* x32 = nullptr;
  uint16_t* x16 = nullptr;
  double* x64 = nullptr;
  uint32_t* bnd = nullptr; // certificate maxima per operand copy (PackArgs::bnd), BND_SLOTS words
  float* rrho = nullptr;   // per-row rounding radius ||x~ - x|| of the FP32 (slot 0) and BF16
                           // (slot 1) operand rows, rounded up: [2][rows_pad] (PackArgs::rrho)
  uint8_t* x8 = nullptr;   // FP8 (E4M3) rows with their scale, rows_pad x ld8 bytes
  int dp8 = 0;             // FP8 row bytes (d rounded up to 64, 128 or 256)
  long ld8 = 0;            // dp8 + 16 (the augmentation unit)
};

namespace knns {

// error reporting: the message of the calling thread's last failure (knns_last_error)
extern thread_local std::string g_err;
knns_status fail(knns_status st, const char* fmt, ...);

inline long round_up(long x, long a) { return (x + a - 1) / a * a; }
void dims_for(int d, int* dp, int* dc);
int fp8_row_bytes(int d);
int default_shortlist(int prec, int k);
// K' of a query: the requested shortlist (at least k), else the built-in rule plus the tuning
// offset (at least k); capped by the searched rows and 2048
int shortlist_for(int prec, int k, int requested, int sl_extra, long m_eff);

struct Round {
  long begin, end, slice_rows;
  int slices;
  int dense;   // every row is a candidate (threshold +inf): unfiltered dense writes
};

// dp: padded FP32 dims of the corpus; dc: the scan's chunk (FP8: the FP8 row bytes)
struct Plan {
  int prec_kernel, qb, dc, dp, groups, n_pad, slices_max, sub_cap, kmax, cmax;
  int carry_cap = 0;   // entries the carry holds per query (cmax, or more with the certificate band)
  std::vector<Round> rounds;
};

// Planner knobs measured by the autotuner (knn-toolbox_amd/tune.py) and persisted as a table.
struct Tuning {
  int min_waves = 5;   // slices: the smallest count filling >= min_waves waves of resident WGs
  int ratio = 0;       // corpus growth between threshold rounds; 0 = x4 below m/16, then x2
  int qb = 0;          // query blocks of 32 per scan wave (tile shape); 0 = the built-in rule
  int sl_extra = 0;    // shortlist offset from the precision's built-in rule (default_shortlist)
};
// Nearest table entry for (precision, d, k, n, m) (log distance over d, k, n, m); defaults
// without a table or without an entry of that precision.
Tuning tuning_for(int prec, int d, int k, long n, long m);

Plan make_plan(long m, int d, long n, int kmax, int cmax, int prec, int cus, int ratio_req,
               int slices_req, int min_waves = 5, bool allow_dense = true, int qb_req = 0,
               bool band = false);

// Relative error bound of the scan's FP32 accumulation of one key (every rounding of it), for
// the certificate: gamma = n u / (1 - n u), n = roundings per key, u = 2^-23 (covers rounding
// modes other than to-nearest), plus the representation error of the norm terms.
double key_gamma(const Plan& P);

// Device buffers of one search.
struct SearchBuffers {
  void* q32 = nullptr;      // packed FP32 query rows, n_pad x ld32
  void* q16 = nullptr;      // packed BF16 query rows (BF16 plans)
  void* q8 = nullptr;       // packed FP8 query rows (FP8 plans)
  void* q64 = nullptr;      // FP64 query master (double corpora)
  void* thr = nullptr;      // u32 per query slot
  void* kc = nullptr;       // shortlist per query slot
  void* kq = nullptr;       // k per query
  void* off = nullptr;      // output offset per query (entries)
  void* self = nullptr;     // self row per query slot (all-kNN), optional
  void* sub = nullptr;      // sub-pools
  void* subcnt = nullptr;   // sub-pool fills
  void* carry[2] = {nullptr, nullptr};
  void* ccnt[2] = {nullptr, nullptr};
  void* flags = nullptr;    // 16 ints: [0] overflow events, [1] slow compactions, [2] uncertified queries
  void* qn2 = nullptr;      // ||q~||^2 per query slot (double)
  void* qrho = nullptr;     // ||q~ - q|| per query slot (double)
  void* odist = nullptr;    // output distances
  void* oidx = nullptr;     // output indices
};

// Byte sizes of the buffers above for a plan (q_slots = query slots of the search).
struct SearchSizes {
  size_t q32, q16, q8, q64, thr, kc, kq, off, self, sub, subcnt, carry, ccnt, flags, odist, oidx, qn2, qrho;
};
SearchSizes search_sizes(const knns_corpus_s* c, const Plan& P, long n, long total_out, bool excl,
                         bool out64);

struct Upload;   // chunked corpus fill (knns_engine.hip)

struct SearchLaunch {
  const knns_corpus_s* c = nullptr;
  const Plan* P = nullptr;
  long n = 0;                  // query slots searched (the real count unless n_dev is given)
  int dtype = KNNS_F32;
  const void* qsrc = nullptr;  // device query rows (dtype), n rows of d values
  bool uniform = true;         // uniform k: the init kernel writes kc / kq / off
  int kuni = 0;
  int cmax = 0;
  bool excl = false;
  long self_offset = 0;
  bool out64 = false;
  int bitonic = 0;
  bool self_preset = false;    // B.self already holds per-query self rows (the init kernel keeps them)
  bool thr_preset = false;     // B.thr already holds the real queries' initial thresholds
  bool certify = true;         // the final kernel checks each query's shortlist certificate
  bool band = false;           // thresholds carry the certificate band (needs certify, not FP8)
  const int* n_dev = nullptr;  // optional device query count <= n (graph-captured searches)
  Upload* feed = nullptr;      // optional: corpus rows still uploading; each scan waits for its rows
};

// Device workspace of a re-search: buffers by slot (the engine's grow-only slots, or a fixed
// array carved in order) and a host-to-device copy of a control block.
struct Workspace {
  virtual ~Workspace() {}
  virtual size_t budget() const = 0;   // bytes the re-search may use
  virtual knns_status get(int slot, size_t bytes, void** out) = 0;
  // enqueues a copy of `bytes` from host memory valid until the caller synchronises `st`
  virtual knns_status upload(void* dst, const void* src, size_t bytes, hipStream_t st) = 0;
};

// Uncertified queries of a finished search (the final kernel left their needed thresholds).
struct ResearchIn {
  const knns_corpus_s* c = nullptr;
  int prec = KNNS_PREC_FP32;
  int dtype = KNNS_F32;
  bool out64 = false, excl = false, bitonic = false;
  long row_begin = 0, row_end = 0;
  const void* qsrc = nullptr;          // device query rows of the search (dtype, d per row)
  std::vector<int> qids;               // query indices to re-search (rows of qsrc)
  std::vector<uint32_t> thr;           // their thresholds (ord): every row below is admitted
  std::vector<int> kq;                 // their k
  std::vector<long> off;               // their output offsets (entries)
  std::vector<int> self;               // their excluded rows (-1: none; used when excl)
  void* odist = nullptr;               // the search's device outputs, rewritten for these queries
  int32_t* oidx = nullptr;
};

// Re-searches R.qids with every row below their threshold admitted (shortlists of up to 2048),
// writes their k results at their offsets and synchronises `st`. Fails when a re-search still
// cannot be certified.
knns_status run_research(const ResearchIn& R, Workspace& ws, hipStream_t st, int cus, long* researched);
// run_research on the engine's own grow-only workspace (serialised with every engine call)
knns_status research_on_engine(const ResearchIn& R, hipStream_t st, long* researched);

// Timing hook: called with 0/1 around each scan launch, 2 after each merge, 3/4 around the final
// kernel (used to record events; nullptr records nothing).
using LaunchHook = std::function<void(int)>;

// Enqueue pack -> init -> rounds -> final on `st`. No host synchronisation, no memset nodes:
// the sequence can be captured into a graph.
hipError_t enqueue_search(const SearchLaunch& L, const SearchBuffers& B, hipStream_t st,
                          const LaunchHook* hook);

}  // namespace knns

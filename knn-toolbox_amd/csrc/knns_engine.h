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

struct Round {
  long begin, end, slice_rows;
  int slices;
  int dense;   // every row is a candidate (threshold +inf): unfiltered dense writes
};

// dp: padded FP32 dims of the corpus; dc: the scan's chunk (FP8: the FP8 row bytes)
struct Plan {
  int prec_kernel, qb, dc, dp, groups, n_pad, slices_max, sub_cap, kmax, cmax;
  std::vector<Round> rounds;
};

// Planner knobs measured by the autotuner (knn-toolbox_amd/tune.py) and persisted as a table.
struct Tuning {
  int min_waves = 5;   // slices: the smallest count filling >= min_waves waves of resident WGs
  int ratio = 0;       // corpus growth between threshold rounds; 0 = x4 below m/16, then x2
};
// Nearest table entry for (precision, d, k, n) (log distance over d, k, n); defaults without a
// table or without an entry of that precision.
Tuning tuning_for(int prec, int d, int k, long n);

Plan make_plan(long m, int d, long n, int kmax, int cmax, int prec, int cus, int ratio_req,
               int slices_req, int min_waves = 5);

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
  void* flags = nullptr;    // 16 ints
  void* odist = nullptr;    // output distances
  void* oidx = nullptr;     // output indices
};

// Byte sizes of the buffers above for a plan (q_slots = query slots of the search).
struct SearchSizes {
  size_t q32, q16, q8, q64, thr, kc, kq, off, self, sub, subcnt, carry, ccnt, flags, odist, oidx;
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
  const int* n_dev = nullptr;  // optional device query count <= n (graph-captured searches)
  Upload* feed = nullptr;      // optional: corpus rows still uploading; each scan waits for its rows
};

// Timing hook: called with 0/1 around each scan launch, 2 after each merge, 3/4 around the final
// kernel (used to record events; nullptr records nothing).
using LaunchHook = std::function<void(int)>;

// Enqueue pack -> init -> rounds -> final on `st`. No host synchronisation, no memset nodes:
// the sequence can be captured into a graph.
hipError_t enqueue_search(const SearchLaunch& L, const SearchBuffers& B, hipStream_t st,
                          const LaunchHook* hook);

}  // namespace knns

// Internal declarations shared by the HIP kernels and the host engine (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knns {

// Precision of the distance GEMM in the scan kernel.
enum Prec : int { PREC_F32 = 0, PREC_BF16 = 1, PREC_FP8 = 2 };

// Candidate entries are 64-bit keys: high word = order-preserving image of the float key
// s = ||r||^2/2 - q.r, low word = corpus row. Unsigned comparison of two entries is the
// lexicographic (key, index) order, which is the tie rule of demo/knn.m:5 (stable sort).
__host__ __device__ inline uint32_t f2ord(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t v) {
  uint32_t u = (v & 0x80000000u) ? (v & 0x7fffffffu) : ~v;
  return __builtin_bit_cast(float, u);
}
constexpr uint32_t ORD_POS_INF = 0xff800000u;  // f2ord(+inf)
constexpr uint32_t ORD_NEG_INF = 0x007fffffu;  // f2ord(-inf)

// Packed row layouts (elements per row). Main dims are padded with zeros to DP; the augmentation
// block carries -||r||^2/2 for corpus rows and ones for query rows so that the MFMA accumulates
// q.r - ||r||^2/2 directly.
//   FP32: [DP main][hi, lo, 0, 0]           lane half h reads aug element h
//   BF16: [DP main][8 aug for half 0][8 zeros for half 1]; corpus aug = 3-term bf16 split
inline int aug_f32() { return 4; }
inline int aug_bf16() { return 16; }

// Certificate band of the shortlist thresholds. With band on, a threshold computed from a K'-th
// approximate key tau is f(tau) = tau + D(tau) rho + rho^2 / 2 + 2 alpha, where D(tau) = sqrt(2 tau +
// ||q~||^2) is the approximate distance of a key-tau row, rho = ||q~ - q|| + max ||r~ - r|| and alpha
// the accumulation bound: every row within the error band of the K'-th row is kept, so the final
// certificate (k-th exact distance vs the last threshold) holds without a fixed shortlist margin.
// f is increasing in tau, so thresholds stay non-increasing over rounds.
struct BandArgs {
  int on;
  const double* qn2;     // ||q~||^2 per query slot
  const double* qrho;    // ||q~ - q|| per query slot
  const uint32_t* bnd;   // corpus maxima (PackArgs::bnd)
  int bslot;             // operand slot in bnd
  double gamma;          // key_gamma of the plan
};

struct ScanArgs {
  const void* X;         // packed corpus rows
  const void* Qp;        // packed query rows
  long ldx, ldq;         // row strides in elements
  int dp;                // padded main dims (multiple of the chunk)
  long row_begin, row_end;  // this round's corpus range
  long slice_rows;       // rows per slice (multiple of 64)
  int slices;            // slices (domains per query) in this round
  int slices_max;        // stride of the sub-pool array
  int n_pad;             // padded query count
  const uint32_t* thr;   // per query ordered threshold (strict): admit f2ord(s) < thr
  const int* kc;         // per query shortlist size
  uint64_t* sub;         // sub-pools [q][slices_max][2][sub_cap]
  int* sub_cnt;          // [q][slices_max][2]
  int sub_cap;
  const int* self_ids;   // optional: corpus row excluded for each query (all-kNN), -1 = none
  int* flags;            // [0] = overflow counter, [1] = slow-path compactions
  int dense;             // 1: no filter; every row is written at its slice-relative position of the
                         // (query, slice, half 0) sub-pool (slice_rows <= sub_cap), fills = rows
  const int* n_dev;      // optional device query count: workgroups whose queries all lie past it
                         // return at once (graph-captured searches over a padded slot count)
  BandArgs band;         // in-scan compaction keeps the band above its K'-th entry; a truncation
                         // lowers thr[q] (atomicMin) so the merge never raises it past dropped rows
  uint32_t* thr_w;       // thr, writable (the same array)
  // scans: items i = slice * wg_rows + row (i < G = wg_rows x slices) on a 1-D grid of
  // round_up(G, 8) workgroups. The device hands workgroup b to XCD b % 8; workgroup b takes item
  // (b % 8) * ceil(G / 8) + b / 8, so that each XCD runs a contiguous run of items, every row of a
  // slice on one XCD (but at the run's ends), and fetches the slice's rows into its L2 once
  int wg_rows;
};

// Rank of the entry that sets a merge's threshold before the last round, for a shortlist of K
// rows after the fraction `seen` of the search's rows. When the rows are in no particular order,
// the number of the final K nearest rows inside that prefix is Binomial(K, seen); the rank is an
// upper quantile of it, mu + 5.2 sqrt(mu) + 5.4 with mu = K seen, which bounds the binomial tail
// by 1e-7 for every K up to 2048 (checked against scipy in the tests), capped at K. The last
// merge (seen = 1) keeps the K smallest.
__host__ __device__ inline int predicted_rank(int K, double seen) {
  if (seen >= 1.0) return K;
  const double mu = (double)K * (seen > 0.0 ? seen : 0.0);
  const int r = (int)ceil(mu + 5.2 * sqrt(mu) + 5.4);
  return r < K ? r : K;
}

struct MergeArgs {
  const uint64_t* carry_in;  // [q][carry_cap]
  const int* carry_in_cnt;
  uint64_t* carry_out;
  int* carry_out_cnt;
  int carry_cap;
  const uint64_t* sub;
  const int* sub_cnt;
  int sub_cap;
  int slices;            // slices used by the round just scanned
  int slices_max;
  uint32_t* thr;         // updated per query
  const int* kc;
  int n;                 // real queries
  int bitonic;           // 1: exact truncated sort (keeps exactly K); 0: histogram bound
  int wide;              // threshold merge: hold up to 2048 entries per query in registers (else 1024)
  int* flags;
  const int* self_ids;   // optional: corpus row excluded for each query (all-kNN), -1 = none
  const int* n_dev;      // optional: real query count on the device (overrides n)
  BandArgs band;         // keep every entry below f(K'-th key) (up to carry_cap), threshold f(K'-th)
  double seen;           // fraction of the search's rows scanned after this round (1 after the last)
};

struct FinalArgs {
  const uint64_t* carry;
  const int* carry_cnt;
  int carry_cap;
  const int* kc;         // shortlist size per query
  const int* kq;         // k per query
  const long* out_off;   // output offset (in entries) per query
  const void* Xm;        // rerank master corpus (float or double rows)
  long ldxm;
  const void* Qm;        // rerank master queries
  long ldqm;
  int d;
  int master_f64;        // 1: master rows are double
  int out_f64;           // 1: write double distances
  void* out_dist;
  int32_t* out_idx;
  int n;
  const int* n_dev;      // optional: real query count on the device (overrides n)
  // shortlist certificate (certify = 1): a query is certified when every corpus row outside its
  // shortlist provably lies farther than its k-th exact distance. thr[q] holds the final strict
  // threshold (ord of the K'-th approximate key; every row outside the shortlist has a key >= it);
  // the kernel overwrites it with 0 for a certified query, else with the ord of the smallest key
  // threshold whose admitted rows certify it, and counts such queries in flags[2].
  int certify;
  uint32_t* thr;
  int* flags;
  long rows;             // candidate rows of a query (the searched rows, less the excluded self row):
                         // a shortlist shorter than k is complete only when it holds them all
  const double* qn2;     // ||q~||^2 per query (operand as rounded for the scan)
  const double* qrho;    // ||q~ - q|| per query, rounded up
  const uint32_t* bnd;   // corpus maxima (PackArgs::bnd); slot bslot
  int bslot;
  double gamma;          // relative bound of the scan's FP32 accumulation (all roundings of a key)
  // prune = 1 (carries with a certificate band; needs qn2, qrho, bnd): the rerank first ranks the
  // carry's first min(K', cnt) entries exactly, then skips every further entry whose key proves it
  // farther than the k-th of those (the same bound as the certificate's)
  int prune;
  const float* rrho;     // optional: per-row rounding radius of the search's operand rows (rounded
                         // up); the pruning then bounds each band entry with its own row's radius
                         // instead of the corpus maximum
};

struct PackArgs {
  const void* src;       // rows of float or double
  int src_f64;
  long rows;             // real rows
  long rows_pad;         // rows written (padding rows are zero / neutral)
  int d;
  long ld_src;
  float* x32;            // [rows_pad][ld32]
  long ld32;
  uint16_t* x16;         // optional bf16 rows [rows_pad][ld16]
  long ld16;
  double* x64;           // optional double master copy [rows][d]
  int dp;                // padded dims
  int is_query;          // aug = ones instead of -||r||^2/2
  uint8_t* x8;           // optional FP8 rows [rows_pad][ld8 bytes] (pack_fp8)
  long ld8;              // bytes per FP8 row: dp8 + 16
  int dp8;               // FP8 row bytes (64, 128 or 256)
  // error-bound inputs of the shortlist certificate (all optional):
  uint32_t* bnd;         // corpus rows: running maxima (float bits, atomicMax) of the rounding radius
                         // ||x~ - x|| and of the norm ||x~|| per operand copy: [BND_RHO + slot],
                         // [BND_NRM + slot], slot = the operand precision (Prec)
  double* qn2;           // query rows: ||q~||^2 of the operand of precision qprec (one per row)
  double* qrho;          // query rows: ||q~ - q|| of that operand, rounded up
  int qprec;             // Prec of the search's operand (which copy qn2 / qrho describe)
  float* rrho;           // corpus rows: ||x~ - x|| per row, rounded up, of the FP32 operand at
                         // rrho[row] and of the BF16 operand at rrho[rrho_ld + row]
  long rrho_ld;
};
constexpr int BND_RHO = 0, BND_NRM = 3, BND_SLOTS = 8;

// Launchers (knns_kernels.hip). All enqueue on `stream` and return hipError_t.
hipError_t launch_pack(const PackArgs& a, hipStream_t stream);
// FP8 rows: [dp8 bytes E4M3 of r / 2^e_r][aug unit: 3 BF16 terms of -||r~||^2 / 2 (r~ = the
// dequantised row), 2 zero bytes, e_r + 127 (E8M0), 7 zero bytes]; query rows carry zero terms.
hipError_t launch_pack_fp8(const PackArgs& a, hipStream_t stream);
hipError_t launch_scan(int prec, int qb, int dc, const ScanArgs& a, int groups, hipStream_t stream);
hipError_t launch_merge(const MergeArgs& a, hipStream_t stream);
hipError_t launch_final(const FinalArgs& a, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t stream);
hipError_t launch_init_queries(uint32_t* thr, int* kc, int* kq, long* off, int* self_ids,
                               int* carry_cnt, long n, long n_pad, int k, int shortlist,
                               long self_offset, long m, const int* n_dev, int write_k,
                               int write_thr, hipStream_t stream);
// dst row i = src row idx[i] (rows of row_bytes, a multiple of 4)
hipError_t launch_gather_rows(void* dst, const void* src, const int* idx, long n, long row_bytes,
                              hipStream_t stream);
hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t stream);
bool scan_supported(int prec, int qb, int dc);
bool scan_qb_ok(int prec, int dp, int dc, int qb);
int scan_wgs_per_cu(int prec, int dp, int dc, int qb);
int scan_stage_rows(int prec, int dp, int dc);

}  // namespace knns

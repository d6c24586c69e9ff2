/* Native C ABI of the MI355X brute-force kNN engine.
 *
 * The drop-in interface (cuknns.h) mirrors the reference. This header is the engine's own
 * interface: device-resident corpora, int32 indices, precision modes, batches with per-query k,
 * all-kNN self exclusion and error codes. It has no counterpart in the reference; the drop-in
 * entry points of cuknns.h are thin wrappers over it. Every function is synchronous on return
 * unless stated otherwise, and thread-safe (calls are serialised per process).
 */
#ifndef KNNS_NATIVE_H
#define KNNS_NATIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  KNNS_OK = 0,
  KNNS_ERR_INVALID = 1,      /* bad argument (shape, k, dtype, precision/corpus mismatch) */
  KNNS_ERR_NOMEM = 2,        /* device or host allocation failed, or over the memory budget */
  KNNS_ERR_DEVICE = 3,       /* HIP runtime error or no device */
  KNNS_ERR_UNSUPPORTED = 4,  /* valid request outside what this library implements */
  KNNS_ERR_OVERFLOW = 5,     /* candidate pool overflow (massive exact ties); results invalid */
  KNNS_ERR_TIMEOUT = 6       /* knns_server_wait: not done within the timeout (the request stays live) */
} knns_status;

typedef enum { KNNS_F32 = 0, KNNS_F64 = 1 } knns_dtype;

typedef enum {
  KNNS_PREC_FP32 = 0,  /* FP32 MFMA shortlist, exact rerank (the drop-in form) */
  KNNS_PREC_BF16 = 1,  /* BF16 MFMA shortlist of `shortlist` candidates, exact rerank */
  KNNS_PREC_FP8 = 2,   /* FP8 e4m3 (per-row scale) shortlist, exact rerank */
  KNNS_PREC_FP64 = 3   /* FP64 corpus/queries: FP32 MFMA shortlist, FP64 rerank + output */
} knns_precision;

typedef enum {
  KNNS_SELECT_AUTO = 0,
  KNNS_SELECT_THRESHOLD = 1, /* heap entry point: threshold filter + insertion, radix-select merge */
  KNNS_SELECT_BITONIC = 2    /* bitonic entry point: truncated bitonic merge of the shortlist */
} knns_select;

/* corpus flags */
#define KNNS_CORPUS_DEVICE_PTR 0x1u /* `data` is device memory */
#define KNNS_CORPUS_BF16 0x2u       /* also build the BF16 operand copy */
#define KNNS_CORPUS_FP8 0x4u        /* also build the FP8 operand copy */

/* search flags */
#define KNNS_DEVICE_PTRS 0x1u   /* queries and outputs are device memory */
#define KNNS_EXCLUDE_SELF 0x2u  /* query i never returns corpus row self_offset + i */
#define KNNS_DIST_F64 0x4u      /* write double distances (default: float, or double for FP64) */
#define KNNS_KERNEL_TIMES 0x8u  /* record per-kernel device times (knns_stats scan_s / select_s); adds
                                   an event per launch, so leave it off when timing whole searches */
#define KNNS_NO_CERTIFY 0x10u   /* skip the shortlist certificate and the re-search of uncertified
                                   queries: the top-k is then exact only when no true neighbour ranks
                                   past the shortlist by approximate key (diagnostics) */

typedef struct knns_corpus_s* knns_corpus;

typedef struct {
  int precision;     /* knns_precision */
  int select;        /* knns_select */
  int k;             /* neighbours per query (ignored by knns_search_multi) */
  int shortlist;     /* candidates kept before the rerank; 0 = precision default */
  unsigned flags;    /* KNNS_DEVICE_PTRS | KNNS_EXCLUDE_SELF | KNNS_DIST_F64 | KNNS_KERNEL_TIMES */
  long self_offset;  /* with KNNS_EXCLUDE_SELF: query i is corpus row self_offset + i */
  int ratio;         /* corpus growth factor between threshold rounds; 0 = planner */
  int slices;        /* corpus slices per query; 0 = planner */
  void* stream;      /* hipStream_t to run on; NULL = the engine's stream */
  int qb;            /* query blocks of 32 per scan wave (tile shape: 1, 2; 3, 4 for BF16 with
                        d <= 128); 0 = planner / tuning table; unsupported values fall back */
} knns_params;

typedef struct {
  double total_s;    /* whole call, device time */
  double scan_s;     /* fused distance + filter kernels (with KNNS_KERNEL_TIMES, else 0) */
  double select_s;   /* merge + final (rerank) kernels (with KNNS_KERNEL_TIMES, else 0) */
  long rounds;
  long slow_compactions;
  long candidates;   /* total entries appended by the scan kernels */
  int groups, slices_max, qb, dc, dp;
  /* Shortlist certificate: a query is certified when every corpus row outside its shortlist
   * provably lies farther than its k-th exact distance (bounds: operand rounding radii and the
   * FP32 accumulation error of the approximate keys). Uncertified queries of the exact forms
   * (FP32, BF16, FP64) are searched again with every row inside the bound admitted, so their
   * results are exact by construction (a BF16 query whose bound admits more rows than the
   * re-search's 2048-row shortlist is searched once more on the FP32 rows); FP8 only counts
   * them. */
  long uncertified;  /* queries whose first shortlist was not certified */
  long researched;   /* queries searched again */
  double research_s; /* device time of the re-search (with KNNS_KERNEL_TIMES, else 0) */
  long shortlist;    /* K' of the search (largest over its queries) */
} knns_stats;

/* fills *p with the defaults of a precision */
void knns_params_init(knns_params* p, int precision, int k);

knns_status knns_corpus_create(knns_corpus* out, const void* data, int dtype, long m, int d,
                               unsigned flags);
knns_status knns_corpus_destroy(knns_corpus c);
long knns_corpus_rows(knns_corpus c);
int knns_corpus_dim(knns_corpus c);

/* n queries (row-major n x d of `dtype`), k = p->k. dist: n*k (float, or double for FP64 or
 * KNNS_DIST_F64), idx: n*k int32, row i = neighbours of query i sorted by (distance, index). */
knns_status knns_search(knns_corpus c, const void* queries, int dtype, long n,
                        const knns_params* p, void* dist, int32_t* idx);

/* Packed batch with per-query k: query i writes kq[i] entries at offset sum_{j<i} kq[j]. */
knns_status knns_search_multi(knns_corpus c, const void* queries, int dtype, long n,
                              const int* kq, const knns_params* p, void* dist, int32_t* idx);

/* nb batches of independent requests packed into one search (one launch sequence): batch b has
 * sizes[b] consecutive query rows, k = ks[b] (ks NULL: p->k for every batch) and, when self_rows
 * is given and self_rows[b] >= 0, query j of batch b never returns corpus row self_rows[b] + j
 * (all-kNN of a row range; -1: no exclusion for that batch). Outputs are packed per query as in
 * knns_search_multi. No counterpart in the reference, whose C path searches one query set per
 * call (src/cuknnBitonic.cu:422); it is the uneven-batch form of the all-kNN demo
 * (demo/allknnTest.m:30-31), with self dropped by index. */
knns_status knns_search_batches(knns_corpus c, const void* queries, int dtype, long nb,
                                const long* sizes, const int* ks, const long* self_rows,
                                const knns_params* p, void* dist, int32_t* idx);

/* knns_search restricted to corpus rows [row_begin, row_end) (0 <= row_begin < row_end <= rows);
 * indices are corpus rows. No counterpart in the reference's exported API; it is the search of
 * one bucket of gpuknnLSH (cuknns.h:79) over a corpus uploaded once for all buckets. */
knns_status knns_search_rows(knns_corpus c, long row_begin, long row_end, const void* queries,
                             int dtype, long n, const knns_params* p, void* dist, int32_t* idx);

/* knns_search without any device or pinned-memory allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Queries and outputs are device pointers (p->flags must hold
 * KNNS_DEVICE_PTRS), k is uniform (p->k), p->stream is used (NULL: the null stream), and the call
 * returns after the search completed; KNNS_KERNEL_TIMES is rejected. The workspace comes from a
 * fixed device array of KNNS_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends; the planner lowers the scan's slice count until the workspace fits (results
 * are the same, only the scan's parallelism changes) and a search that does not fit even with one
 * slice fails with KNNS_ERR_NOMEM. Calls in one process are serialised; the array keeps the last
 * call's contents and is shared by every call on that GPU, so searches issued on one GPU from
 * outside this serialisation (another copy of the library in the process) would race. */
#define KNNS_NOALLOC_SCRATCH_BYTES (256L << 20)
knns_status knns_search_noalloc(knns_corpus c, const void* queries, int dtype, long n,
                                const knns_params* p, void* dist, int32_t* idx);

/* knns_corpus_create + knns_search + knns_corpus_destroy in one call, for a corpus searched once
 * (the whole-problem call of gpuknnsBitonic / gpuknnsHeap, src/cuknnBitonic.cu:422-618): host rows
 * (m x d of `dtype`) upload on a second stream while the first threshold rounds, which scan a
 * growing prefix of the corpus, already run. Host queries and outputs (KNNS_DEVICE_PTRS is
 * rejected); any precision (a BF16 or FP8 search packs its operand copy with the FP32 rows as
 * they arrive); results as knns_search. */
knns_status knns_search_once(const void* data, int dtype, long m, int d, const void* queries,
                             long n, const knns_params* p, void* dist, int32_t* idx);

/* Planner view: the launch configuration knns_search would use (no GPU work). */
knns_status knns_plan_describe(long m, int d, long n, int k, int precision, knns_stats* out);

/* Rank of the entry that sets a certified search's threshold after the fraction `seen` of its
 * rows, for a shortlist of `shortlist` rows (the shortlist itself at seen = 1). No counterpart in
 * the reference; exposed so that the bound it is built on can be checked. */
int knns_predicted_rank(int shortlist, double seen);

const char* knns_last_error(void);
knns_status knns_last_stats(knns_stats* out);
int knns_device_count(void);
/* releases the cached workspace of the engine */
void knns_release_workspace(void);
/* Replaces the planner's tuning table (text: "precision d k n m min_waves ratio qb sl_extra [ms]" per line,
 * as written by knn-toolbox_amd/tune.py); NULL or "" clears it (built-in defaults). Without a
 * call the table is read from $KNNS_TUNING_TABLE or <library dir>/../tuning/gfx950.tsv. */
knns_status knns_set_tuning_table(const char* path);

/* ---- serving: continuous batching over a resident corpus ----
 * No counterpart in the reference (its API is one synchronous call per query batch); this is
 * the serving form of the same search for callers that submit many independent requests.
 * Requests are split into pieces packed, in arrival order, into batches of query slots. The
 * server runs `lanes` batches at once (default 2), each lane with its own device workspace, stream
 * and captured HIP graph per batch-capacity bucket (64, 256, 1024, ... up to max_batch slots): a
 * dispatcher thread packs, uploads and launches the next batch on a free lane while the previous
 * one runs, so their kernels overlap on the device, and a completer thread retires batches in
 * order. Work that does not fit the free lanes waits in the queue (deferred, never refused); a
 * request larger than a batch is served in several. A request's result does not depend on what
 * it is batched with: it equals knns_search on its own (shortlists certified, as there). */
typedef struct knns_server_s* knns_server;

typedef struct {
  int precision;       /* knns_precision (FP32, BF16, FP8; FP64 for double corpora) */
  int max_k;           /* largest k a request may ask for, 1..1024 */
  long max_batch;      /* query slots of the largest bucket (>= 1) */
  long mem_budget;     /* device bytes the server may hold; 0 = no limit. Buckets whose workspace
                          (times the lanes) does not fit are dropped at creation (the largest first) */
  int lanes;           /* batches in flight at once, 1 or 2; 0 = 2 */
} knns_server_config;

typedef struct {
  long requests, queries, batches;
  double p50_ms, p99_ms, max_ms;  /* request latency: submit to results written */
  double mean_batch;              /* queries per batch */
  double busy_s;                  /* wall time from the first launch to the last retired batch */
  long buckets, max_batch;        /* after the memory budget */
  long workspace_bytes;           /* all lanes */
  long lanes;
  long overlapped;                /* batches whose device work began before the previous one ended */
  double device_s;                /* summed device time of the batches (upload to download) */
  long researched;                /* queries searched again by the shortlist certificate */
} knns_server_stats;

knns_status knns_server_create(knns_server* out, knns_corpus c, const knns_server_config* cfg);
/* Copies the n queries (host memory, dtype of the corpus); results are written to dist (n*k,
 * float, or double for double corpora) and idx (n*k int32), which must stay valid until
 * knns_server_wait returns for the ticket. n may exceed max_batch (the request is split). */
knns_status knns_server_submit(knns_server s, const void* queries, int dtype, long n, int k,
                               void* dist, int32_t* idx, unsigned long long* ticket);
/* Blocks until the request is done (timeout_s < 0: no limit) and returns its status; a ticket
 * can be waited for once it is done. KNNS_ERR_INVALID for an unknown ticket, KNNS_ERR_TIMEOUT
 * when the timeout runs out first: the request stays live, its output buffers must stay valid,
 * and the ticket can be waited for again. */
knns_status knns_server_wait(knns_server s, unsigned long long ticket, double timeout_s);
knns_status knns_server_stats_get(knns_server s, knns_server_stats* out, int reset);
/* Finishes the requests already submitted, then frees the server. */
knns_status knns_server_destroy(knns_server s);

#ifdef __cplusplus
}
#endif
#endif /* KNNS_NATIVE_H */

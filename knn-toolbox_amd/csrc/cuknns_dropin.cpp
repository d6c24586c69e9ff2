// Drop-in implementation of include/cuknns.h over the native engine (include/knns_native.h).
// Compiled twice: without __DOUBLE__ into libcuknn.so and with it into libcuknn_double.so, like
// the reference's libcuknn.a / libcuknn_double.a (Makefile:57,64).
#include <float.h>
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../include/cuknns.h"
#include "../../include/knns_native.h"

namespace {

#ifdef __DOUBLE__
constexpr int kDtype = KNNS_F64;
constexpr int kPrec = KNNS_PREC_FP64;
#else
constexpr int kDtype = KNNS_F32;
constexpr int kPrec = KNNS_PREC_FP32;
#endif

thread_local std::string g_err;
thread_local knntimes g_times = {0.0, 0.0, 0.0};

// Whole-problem driver shared by gpuknnsBitonic (src/cuknnBitonic.cu:422-618) and gpuknnsHeap
// (src/cuknnHeap.cu:394-575): upload the corpus, search, return results as knntype Q x k
// (k leading), elapsed seconds or FLT_MAX on failure (src/cuknnBitonic.cu:581).
double run(knntype* query, knntype* data, knntype* values, knntype* indices, int objects,
           int numQueries, int attributes, int k, int select) {
  g_err.clear();
  g_times = {0.0, 0.0, 0.0};
  if (!query || !data || !values || !indices) {
    g_err = "null pointer argument";
    return FLT_MAX;
  }
  if (objects < 1 || numQueries < 0 || attributes < 1 || k < 1 || k > 1024) {
    char b[160];
    snprintf(b, sizeof(b), "invalid sizes: objects=%d numQueries=%d attributes=%d k=%d", objects,
             numQueries, attributes, k);
    g_err = b;
    return FLT_MAX;
  }
  if (numQueries == 0) return 0.0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) {
    g_err = "hipEventCreate failed (no device?)";
    (void)hipGetLastError();
    return FLT_MAX;
  }
  (void)hipEventRecord(t0, nullptr);
  knns_corpus c = nullptr;
  knns_status st = knns_corpus_create(&c, data, kDtype, objects, attributes, 0);
  std::vector<int32_t> idx;
  knns_stats stats{};
  if (st == KNNS_OK) {
    knns_params p;
    knns_params_init(&p, kPrec, k);
    p.select = select;
    p.flags = KNNS_KERNEL_TIMES;   // knnsGetLastTimes: distance and search stage times
    idx.resize((size_t)numQueries * k);
    st = knns_search(c, query, kDtype, numQueries, &p, values, idx.data());
    knns_last_stats(&stats);
  }
  if (st != KNNS_OK) g_err = knns_last_error();
  knns_corpus_destroy(c);
  (void)hipEventRecord(t1, nullptr);
  (void)hipEventSynchronize(t1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (st != KNNS_OK) return FLT_MAX;
  for (size_t i = 0; i < idx.size(); ++i) indices[i] = (knntype)idx[i];
  g_times.dst_time = stats.scan_s * 1e3;
  g_times.srch_time = stats.select_s * 1e3;
  g_times.knn_time = (double)ms;
  return (double)ms * 1e-3;
}

}  // namespace

extern "C" {

double gpuknnsBitonic(knntype* query, knntype* data, knntype* values, knntype* indices,
                      int objects, int numQueries, int attributes, int k, int numStreams) {
  (void)numStreams;  // one stream; the corpus stays resident for the whole call
  return run(query, data, values, indices, objects, numQueries, attributes, k, KNNS_SELECT_BITONIC);
}

double gpuknnsHeap(knntype* query, knntype* data, knntype* values, knntype* indices, int objects,
                   int numQueries, int attributes, int k, int numStreams) {
  (void)numStreams;
  return run(query, data, values, indices, objects, numQueries, attributes, k,
             KNNS_SELECT_THRESHOLD);
}

// src/knnsplan.cu:25-105 timed both algorithms over streams {1,2,4,8} x 10 iterations on random
// data of the full size. Here the choice comes from a table measured once on MI355X: both entry
// points share the fused scan and return identical results, and the threshold (heap) merge is
// never slower than the truncated bitonic merge for the sizes measured, so plans pick it; streams
// are irrelevant to a resident-corpus engine and stay 1.
void knnsplan(knnplan* plan, long int N, long int Q, long int D, long int k) {
  if (!plan) return;
  plan->pt2Function = &gpuknnsHeap;
  plan->objects = (int)N;
  plan->numQueries = (int)Q;
  plan->dimentions = (int)D;
  plan->k = (int)k;
  plan->numStreams = 1;
}

// src/knnsexecute.cu:6-18
void knnsexecute(knnplan plan, knntype* data, knntype* queries, knntype* KNNdist,
                 knntype* KNNidx) {
  if (!plan.pt2Function) {
    g_err = "plan has no function (knnsplan not called?)";
    return;
  }
  double time = plan.pt2Function(queries, data, KNNdist, KNNidx, plan.objects, plan.numQueries,
                                 plan.dimentions, plan.k, plan.numStreams);
  const char* quiet = getenv("KNNS_QUIET");
  if (!(quiet && quiet[0] == '1')) printf("Time Elapsed: %f\n", time);
}

const char* knnsGetLastError(void) { return g_err.c_str(); }

knntimes knnsGetLastTimes(void) { return g_times; }

}  // extern "C"

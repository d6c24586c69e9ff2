// CPU oracle for the brute-force kNN path. TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and
// only to check results or as the timed CPU baseline. The GPU product path never calls it.
//
// Three restatements live here:
//
//  1. oracle_knnm_f32 / oracle_knnm_f64 -- a line-for-line transcription of the reference's CPU
//     brute force, demo/knn.m:5-9:
//         [D, N] = sort(bsxfun(@plus, (-2)*(Y'*X), dot(Y,Y,1)'));   % knn.m:5
//         N = N(1:k,:);  D = D(1:k,:);                                % knn.m:7-8
//         D = bsxfun(@plus, D, dot(X,X,1));                           % knn.m:9
//     X holds the queries and Y the corpus, one vector per column in MATLAB; here both are
//     row-major with D leading, which is the same memory. MATLAB's sort is stable, so equal keys
//     keep the lower corpus index first. The GEMM is ours: one sequential dot product per
//     (query, corpus) pair, separately rounded multiply and add (this file is compiled with
//     -ffp-contract=off). Indices are returned 0-based; the toolbox adds 1
//     (toolbox/gpuknnBitonic.m:24).
//     This is also the "reference semantics" of the GPU code: the reference ranks corpus points
//     by key = ||r||^2 - 2 q.r (include/utils.h:957-985 builds exactly this matrix) and adds
//     ||q||^2 after selection (include/utils.h:485-491).
//
//  2. oracle_exact_topk -- the FP64 brute force with the canonical distance used everywhere in
//     this project:  dist(q, r) = fold_{i=0..d-1} acc = fma(t_i, t_i, acc),  t_i = double(q_i) -
//     double(r_i), acc starting at +0.0. The GPU rerank kernel evaluates the same expression in
//     the same order, so the two agree bit for bit. Neighbours are ordered by (dist, index)
//     lexicographically; an optional self-exclusion list drops corpus row self[i] for query i
//     (the all-kNN case of demo/allknnTest.m:30-31, done by index instead of by row deletion).
//
//  3. oracle_keys_f32 -- the reference key matrix key[i][j] = fl(fl(-2 * fl(q.r)) + fl(||r||^2))
//     for small inputs, so Python restatements of the reference's selection kernels
//     (src/cuknnHeap.cu, src/cuknnBitonic.cu) can be run on exactly the keys the reference ranks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

template <typename T>
T seq_dot(const T* a, const T* b, int d) {
  T acc = T(0);
  for (int t = 0; t < d; ++t) acc = acc + a[t] * b[t];
  return acc;
}

template <typename T>
void knnm(const T* X, const T* Y, long n, long m, int d, int k, T* Dout, int32_t* Iout) {
  // dot(Y,Y,1)' : corpus norms (knn.m:5)
  std::vector<T> yy(m);
  for (long j = 0; j < m; ++j) yy[j] = seq_dot(Y + j * d, Y + j * d, d);
#pragma omp parallel
  {
    std::vector<T> key(m);
    std::vector<int32_t> perm(m);
#pragma omp for schedule(dynamic, 1)
    for (long i = 0; i < n; ++i) {
      const T* x = X + i * d;
      // (-2)*(Y'*X) + dot(Y,Y,1)'  (one column of the bsxfun result)
      for (long j = 0; j < m; ++j) key[j] = (T(-2) * seq_dot(Y + j * d, x, d)) + yy[j];
      std::iota(perm.begin(), perm.end(), 0);
      // MATLAB sort: ascending and stable
      std::stable_sort(perm.begin(), perm.end(),
                       [&](int32_t a, int32_t b) { return key[a] < key[b]; });
      const T xx = seq_dot(x, x, d);  // dot(X,X,1) (knn.m:9)
      for (int t = 0; t < k; ++t) {
        Dout[i * k + t] = key[perm[t]] + xx;
        Iout[i * k + t] = perm[t];
      }
    }
  }
}

template <typename T>
double canon_dist(const T* q, const T* r, int d) {
  double acc = 0.0;
  for (int t = 0; t < d; ++t) {
    const double diff = double(q[t]) - double(r[t]);
    acc = std::fma(diff, diff, acc);
  }
  return acc;
}

template <typename T>
void exact_topk(const T* Q, const T* R, long n, long m, int d, int k, const int64_t* self,
                double* Dout, int32_t* Iout) {
#pragma omp parallel
  {
    std::vector<std::pair<double, int32_t>> buf(m);
#pragma omp for schedule(dynamic, 1)
    for (long i = 0; i < n; ++i) {
      const T* q = Q + i * d;
      long cnt = 0;
      for (long j = 0; j < m; ++j) {
        if (self && self[i] == j) continue;
        buf[cnt++] = {canon_dist(q, R + j * d, d), int32_t(j)};
      }
      const long kk = std::min<long>(k, cnt);
      std::partial_sort(buf.begin(), buf.begin() + kk, buf.begin() + cnt);
      for (long t = 0; t < k; ++t) {
        if (t < kk) {
          Dout[i * k + t] = buf[t].first;
          Iout[i * k + t] = buf[t].second;
        } else {
          Dout[i * k + t] = HUGE_VAL;
          Iout[i * k + t] = -1;
        }
      }
    }
  }
}

}  // namespace

extern "C" {

void oracle_knnm_f32(const float* X, const float* Y, long n, long m, int d, int k, float* D,
                     int32_t* I) {
  knnm<float>(X, Y, n, m, d, k, D, I);
}

void oracle_knnm_f64(const double* X, const double* Y, long n, long m, int d, int k, double* D,
                     int32_t* I) {
  knnm<double>(X, Y, n, m, d, k, D, I);
}

// dtype: 0 = float32 inputs, 1 = float64 inputs. self may be NULL.
void oracle_exact_topk(const void* Q, const void* R, int dtype, long n, long m, int d, int k,
                       const int64_t* self, double* D, int32_t* I) {
  if (dtype == 0)
    exact_topk<float>((const float*)Q, (const float*)R, n, m, d, k, self, D, I);
  else
    exact_topk<double>((const double*)Q, (const double*)R, n, m, d, k, self, D, I);
}

// Canonical distances for explicit (query, candidate) pairs: D[i*c + t] = dist(Q[i], R[C[i*c+t]]).
void oracle_pair_dist(const void* Q, const void* R, int dtype, long n, int d, int c,
                      const int32_t* C, double* D) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i)
    for (int t = 0; t < c; ++t) {
      const long j = C[i * c + t];
      D[i * c + t] = dtype == 0 ? canon_dist((const float*)Q + i * d, (const float*)R + j * d, d)
                                : canon_dist((const double*)Q + i * d, (const double*)R + j * d, d);
    }
}

// key[i*m + j] = fl(fl(-2*fl(q_i.r_j)) + fl(||r_j||^2)), sequential FP32 sums.
void oracle_keys_f32(const float* Q, const float* R, long n, long m, int d, float* key) {
  std::vector<float> yy(m);
  for (long j = 0; j < m; ++j) yy[j] = seq_dot(R + j * d, R + j * d, d);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < m; ++j) key[i * m + j] = (-2.0f * seq_dot(R + j * d, Q + i * d, d)) + yy[j];
}

}  // extern "C"

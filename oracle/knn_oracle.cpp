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
#include <type_traits>
#include <vector>

namespace {

template <typename T>
T seq_dot(const T* a, const T* b, int d) {
  T acc = T(0);
  for (int t = 0; t < d; ++t) acc = acc + a[t] * b[t];
  return acc;
}

// Order-preserving unsigned image of a key (equal keys -> equal images: -0 is folded onto +0).
inline uint32_t key_bits(float f) {
  if (f == 0.0f) f = 0.0f;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline uint64_t key_bits(double f) {
  if (f == 0.0) f = 0.0;
  uint64_t u;
  std::memcpy(&u, &f, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

// Stable ascending sort of perm[0..m) by key[perm[i]]: LSD radix sort over the order-preserving
// key images, 8 bits per pass. Every pass is stable, so equal keys keep their input order -- the
// result is the permutation std::stable_sort(perm, key[a] < key[b]) gives (MATLAB's sort).
template <typename T>
void stable_key_sort(const T* key, int32_t* perm, int32_t* tmp, long m) {
  using U = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
  std::vector<U> kb(m), kt(m);
  for (long j = 0; j < m; ++j) kb[j] = key_bits(key[perm[j]]);
  for (unsigned sh = 0; sh < 8 * sizeof(U); sh += 8) {
    long cnt[257] = {0};
    for (long j = 0; j < m; ++j) ++cnt[((kb[j] >> sh) & 0xff) + 1];
    if (cnt[1 + ((kb[0] >> sh) & 0xff)] == m) continue;   // one bucket: the pass is the identity
    for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for (long j = 0; j < m; ++j) {
      const long o = cnt[(kb[j] >> sh) & 0xff]++;
      tmp[o] = perm[j];
      kt[o] = kb[j];
    }
    std::memcpy(perm, tmp, m * sizeof(int32_t));
    kb.swap(kt);
  }
}

template <typename T>
void knnm(const T* X, const T* Y, long n, long m, int d, int k, T* Dout, int32_t* Iout) {
  // dot(Y,Y,1)' : corpus norms (knn.m:5)
  std::vector<T> yy(m);
  for (long j = 0; j < m; ++j) yy[j] = seq_dot(Y + j * d, Y + j * d, d);
  // The GEMM is blocked (QB queries x JT corpus rows per tile, the tile transposed once per query
  // block) but every dot product keeps its own sequential order over the d terms, so each key is
  // bit-identical to seq_dot's: the blocking only vectorises across independent dot products.
  constexpr int QB = 8, JT = 64;
#pragma omp parallel
  {
    std::vector<T> key((size_t)QB * m);
    std::vector<int32_t> perm(m), tmp(m);
    std::vector<T> yt((size_t)d * JT);
#pragma omp for schedule(dynamic, 1)
    for (long i0 = 0; i0 < n; i0 += QB) {
      const int nq = (int)std::min<long>(QB, n - i0);
      for (long j0 = 0; j0 < m; j0 += JT) {
        const int nj = (int)std::min<long>(JT, m - j0);
        for (int jj = 0; jj < JT; ++jj)
          for (int t = 0; t < d; ++t) yt[(size_t)t * JT + jj] = jj < nj ? Y[(j0 + jj) * d + t] : T(0);
        for (int qi = 0; qi < nq; ++qi) {
          const T* x = X + (i0 + qi) * d;
          T acc[JT];
          for (int jj = 0; jj < JT; ++jj) acc[jj] = T(0);
          for (int t = 0; t < d; ++t) {
            const T xt = x[t];
            const T* yr = &yt[(size_t)t * JT];
            for (int jj = 0; jj < JT; ++jj) acc[jj] = acc[jj] + yr[jj] * xt;
          }
          // (-2)*(Y'*X) + dot(Y,Y,1)'  (one column of the bsxfun result)
          for (int jj = 0; jj < nj; ++jj) key[(size_t)qi * m + j0 + jj] = (T(-2) * acc[jj]) + yy[j0 + jj];
        }
      }
      for (int qi = 0; qi < nq; ++qi) {
        const long i = i0 + qi;
        const T* kq = &key[(size_t)qi * m];
        std::iota(perm.begin(), perm.end(), 0);
        stable_key_sort(kq, perm.data(), tmp.data(), m);   // MATLAB sort: ascending and stable
        const T xx = seq_dot(X + i * d, X + i * d, d);     // dot(X,X,1) (knn.m:9)
        for (int t = 0; t < k; ++t) {
          Dout[i * k + t] = kq[perm[t]] + xx;
          Iout[i * k + t] = perm[t];
        }
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
      // four rows at a time: four independent FMA chains, each in the canonical order (the
      // results are those of canon_dist; only the latency of the chains overlaps)
      long j = 0;
      for (; j + 4 <= m; j += 4) {
        const T* r0 = R + j * d;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int t = 0; t < d; ++t) {
          const double qt = double(q[t]);
          const double t0 = qt - double(r0[t]), t1 = qt - double(r0[d + t]);
          const double t2 = qt - double(r0[2 * d + t]), t3 = qt - double(r0[3 * d + t]);
          a0 = std::fma(t0, t0, a0);
          a1 = std::fma(t1, t1, a1);
          a2 = std::fma(t2, t2, a2);
          a3 = std::fma(t3, t3, a3);
        }
        const double a[4] = {a0, a1, a2, a3};
        for (int u = 0; u < 4; ++u)
          if (!(self && self[i] == j + u)) buf[cnt++] = {a[u], int32_t(j + u)};
      }
      for (; j < m; ++j) {
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

// Canonical distances for explicit (query, candidate) pairs: D[i*c + t] = dist(Q[i], R[C[i*c+t]]),
// Q and R both of `dtype`, R holding m rows. A negative index (the -1 of an empty slot past a
// bucket smaller than k) gives +inf; an index >= m is never read and gives NaN.
void oracle_pair_dist(const void* Q, const void* R, int dtype, long n, long m, int d, int c,
                      const int32_t* C, double* D) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i)
    for (int t = 0; t < c; ++t) {
      const long j = C[i * c + t];
      if (j < 0) { D[i * c + t] = HUGE_VAL; continue; }
      if (j >= m) { D[i * c + t] = std::nan(""); continue; }
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

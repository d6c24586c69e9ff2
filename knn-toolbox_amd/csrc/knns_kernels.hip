// HIP kernels for the brute-force kNN path on CDNA4 (gfx950).
//
//   pack_kernel   rows -> padded FP32 / BF16 operand rows with the norm augmentation, FP64 master
//   scan_kernel   fused MFMA distance + per-query threshold filter + append (one wave per
//                 (query group, corpus slice)); replaces the GEMM + norm kernels + selection of the
//                 reference (include/utils.h:883-985, src/cuknnBitonic.cu:45-254,
//                 src/cuknnHeap.cu:146-291) without ever writing the n x m key matrix
//   merge_kernel  per query: union of the round's candidates and the carried shortlist, exact
//                 K-th entry by radix select (threshold mode) or truncated bitonic sort (bitonic
//                 mode), new strict threshold for the next round
//   final_kernel  per query: sort the shortlist, exact rerank with the canonical FP64 distance,
//                 order by (distance, index), write the k nearest
//
// Keys. The scan ranks s = ||r~||^2/2 - q~.r~ where q~, r~ are the operands as rounded for the
// MFMA precision and ||r~||^2 is the exact norm of the rounded row. The reference ranks
// ||r||^2 - 2 q.r (include/utils.h:957-985), i.e. 2s. Using the norm of the rounded row makes s
// exactly (||q~ - r~||^2 - ||q~||^2)/2, so the shortlist error scales with ||q - r||, not ||q||.
#include <cstdlib>
#include <type_traits>

#include "knns_internal.h"

namespace knns {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint64_t w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint64_t w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t w = __shfl_up(v, o, 64);
    if (lane >= o) v += w;
  }
  return v;
}
__device__ __forceinline__ int bits64(uint64_t x) { return x ? 64 - __clzll(x) : 0; }

// The lane index through an empty asm. Rarely taken paths inside the scan loop (appends, in-scan
// compaction) use it so that the values they derive from the lane index are computed there,
// instead of being hoisted out of the loop into registers that are then held for the whole scan.
__device__ __forceinline__ int lane_opaque() {
  int l = (int)(threadIdx.x & 63);
  asm volatile("" : "+v"(l));
  return l;
}

__device__ __forceinline__ double next_up_f32(double x) {   // a float strictly above x
  const float f = (float)x;
  float g = (double)f > x ? f : __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) + (f >= 0.f ? 1u : -1u));
  if (g == 0.f) g = 0x1p-149f;
  return (double)g;
}

// Threshold of the certificate band (BandArgs) for a K'-th key of order image t: the order image
// of a float strictly above f(tau) = tau + D(tau) rho + rho^2 / 2 + 2 alpha; never below t.
__device__ __forceinline__ uint32_t band_thr(const BandArgs& b, int q, uint32_t t) {
  if (t >= ORD_POS_INF) return t;
  const double tau = (double)ord2f(t);
  const double qn2 = b.qn2[q];
  const double rmax = (double)__builtin_bit_cast(float, b.bnd[BND_RHO + b.bslot]);
  const double nmax = (double)__builtin_bit_cast(float, b.bnd[BND_NRM + b.bslot]);
  const double alpha = b.gamma * (__builtin_sqrt(qn2) * nmax + 0.5 * nmax * nmax);
  const double rho = b.qrho[q] + rmax;
  const double D = __builtin_sqrt(__builtin_fmax(0.0, 2.0 * tau + qn2));
  const uint32_t r = f2ord((float)next_up_f32(tau + D * rho + 0.5 * rho * rho + 2.0 * alpha));
  return r > t ? r : t;
}

// ------------------------------------------------------------------------------------------------
// K-th smallest of N distinct 64-bit entries (radix select, 8 bits per level), by one wave
// (WAVE) or by the whole block (blockDim multiple of 64, <= 256). get(i) returns entry i. Requires 1 <= K <= N.
// lds: >= 256 + 16 uint32 of shared scratch.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool WAVE>
__device__ __forceinline__ void group_sync() {
  if constexpr (WAVE) wave_sync(); else __syncthreads();
}

// each(f) calls f(e) for every entry this thread holds (the entries of the group are the union).
template <bool WAVE, class Each>
__device__ __forceinline__ uint64_t group_kth_each(Each each, int N, int K, uint32_t* lds) {
  uint32_t* hist = lds;                          // 256 bins
  uint64_t* red = (uint64_t*)(lds + 256);        // 2 x (waves) u64, waves <= 4
  const int tid = WAVE ? lane_opaque() : (int)threadIdx.x;
  const int nt = WAVE ? 64 : (int)blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  uint64_t mn = ~0ull, mx = 0;
  each([&](uint64_t e) {
    mn = e < mn ? e : mn;
    mx = e > mx ? e : mx;
  });
  mn = wave_min_u64(mn);
  mx = wave_max_u64(mx);
  if (lane == 0) { red[wv] = mn; red[4 + wv] = mx; }
  group_sync<WAVE>();
  mn = red[0]; mx = red[4];
  for (int w = 1; w < nw; ++w) {
    mn = red[w] < mn ? red[w] : mn;
    mx = red[4 + w] > mx ? red[4 + w] : mx;
  }
  group_sync<WAVE>();
  if (K >= N) return mx;
  if (K <= 1) return mn;
  uint64_t base = mn;
  int below = 0;
  int sh = bits64(mx - mn);
  sh = sh > 8 ? sh - 8 : 0;
  uint64_t result = mx;
  for (int level = 0; level < 9; ++level) {
    for (int b = tid; b < 256; b += nt) hist[b] = 0;
    group_sync<WAVE>();
    each([&](uint64_t e) {
      if (e >= base) {
        uint64_t r = (e - base) >> sh;
        if (r < 256) atomicAdd(&hist[(int)r], 1u);
      }
    });
    group_sync<WAVE>();
    // wave 0 finds the bin holding the K-th entry
    if (wv == 0) {
      uint32_t c0 = hist[lane * 4 + 0], c1 = hist[lane * 4 + 1], c2 = hist[lane * 4 + 2],
               c3 = hist[lane * 4 + 3];
      uint32_t tot = c0 + c1 + c2 + c3;
      uint32_t incl = wave_incl_scan_u32(tot, lane);
      uint32_t excl = incl - tot;
      const uint32_t need = (uint32_t)(K - below);
      // first lane whose inclusive count reaches `need`
      uint64_t hit = __builtin_amdgcn_ballot_w64(incl >= need);
      int fl = __ffsll((long long)hit) - 1;
      if (lane == fl) {
        uint32_t acc = excl;
        int b = 0;
        uint32_t cb = c0;
        if (acc + c0 >= need) { b = 0; cb = c0; }
        else if (acc + c0 + c1 >= need) { b = 1; acc += c0; cb = c1; }
        else if (acc + c0 + c1 + c2 >= need) { b = 2; acc += c0 + c1; cb = c2; }
        else { b = 3; acc += c0 + c1 + c2; cb = c3; }
        red[0] = (uint64_t)(lane * 4 + b);
        red[1] = (uint64_t)acc;      // entries below the chosen bin (this level)
        red[2] = (uint64_t)cb;       // entries in the chosen bin
      }
    }
    group_sync<WAVE>();
    const int bstar = (int)red[0];
    const int blw = (int)red[1];
    const int cbin = (int)red[2];
    group_sync<WAVE>();
    below += blw;
    base += (uint64_t)bstar << sh;
    if (sh == 0) { result = base; break; }
    if (below + cbin == K) {
      // every entry of the bin is kept: the K-th entry is the bin's maximum
      const uint64_t hi = base + ((1ull << sh) - 1);
      uint64_t m2 = 0;
      each([&](uint64_t e) {
        if (e >= base && e <= hi) m2 = e > m2 ? e : m2;
      });
      m2 = wave_max_u64(m2);
      if (lane == 0) red[4 + wv] = m2;
      group_sync<WAVE>();
      m2 = red[4];
      for (int w = 1; w < nw; ++w) m2 = red[4 + w] > m2 ? red[4 + w] : m2;
      group_sync<WAVE>();
      result = m2;
      break;
    }
    sh = sh > 8 ? sh - 8 : 0;
  }
  return result;
}

template <bool WAVE, class Get>
__device__ __forceinline__ uint64_t group_kth(Get get, int N, int K, uint32_t* lds) {
  const int tid = WAVE ? lane_opaque() : (int)threadIdx.x;
  const int nt = WAVE ? 64 : (int)blockDim.x;
  return group_kth_each<WAVE>([&](auto f) { for (int i = tid; i < N; i += nt) f(get(i)); }, N, K, lds);
}

// smallest float >= x (x >= 0, finite)
__device__ __forceinline__ float f32_up(double x) {
  const float f = (float)x;
  return (double)f >= x ? f : __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) + 1u);
}
// sqrt(x) rounded up to float (the double sqrt is within half an ulp; the factor covers it)
__device__ __forceinline__ float sqrt_up(double x) { return f32_up(__builtin_sqrt(x) * (1.0 + 0x1p-50)); }

// Certificate inputs of a packed row: corpus rows raise the running maxima of the rounding radius
// ||x~ - x|| and of ||x~|| (float bits of non-negative values order like unsigned integers); a
// query row records ||q~||^2 and ||q~ - q|| of the search's operand.
__device__ __forceinline__ void pack_bounds(const PackArgs& a, long row, int slot, double nrm2, double err2) {
  if (a.is_query) {
    if (a.qn2 && a.qprec == slot) {
      a.qn2[row] = nrm2;
      a.qrho[row] = (double)sqrt_up(err2);
    }
  } else if (a.bnd && row < a.rows) {
    const float rho = sqrt_up(err2);
    if (a.rrho && slot <= PREC_BF16) a.rrho[slot * a.rrho_ld + row] = rho;
    // an atomic only when the row raises the running maximum (the maxima settle after a few rows;
    // a stale read only costs an unneeded atomic)
    const uint32_t rb = __builtin_bit_cast(uint32_t, rho), nb = __builtin_bit_cast(uint32_t, sqrt_up(nrm2));
    if (rb > __hip_atomic_load(&a.bnd[BND_RHO + slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&a.bnd[BND_RHO + slot], rb);
    if (nb > __hip_atomic_load(&a.bnd[BND_NRM + slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&a.bnd[BND_NRM + slot], nb);
  }
}

// ------------------------------------------------------------------------------------------------
// pack: one wave per row
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_kernel(PackArgs a) {
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= a.rows_pad) return;
  const bool real = row < a.rows;
  double n32 = 0.0, n16 = 0.0, e32 = 0.0, e16 = 0.0;
  for (int j = lane; j < a.dp; j += 64) {
    double v = 0.0;
    if (real && j < a.d) {
      v = a.src_f64 ? ((const double*)a.src)[row * a.ld_src + j]
                    : (double)((const float*)a.src)[row * a.ld_src + j];
      if (a.x64) a.x64[row * a.d + j] = v;
    }
    const float f = (float)v;
    a.x32[row * a.ld32 + j] = f;
    n32 += (double)f * (double)f;
    e32 += ((double)f - v) * ((double)f - v);
    if (a.x16) {
      const uint16_t b = f2bf(f);
      a.x16[row * a.ld16 + j] = b;
      const double bd = (double)bf2f(b);
      n16 += bd * bd;
      e16 += (bd - v) * (bd - v);
    }
  }
  n32 = wave_sum_f64(n32);
  n16 = wave_sum_f64(n16);
  e32 = wave_sum_f64(e32);
  e16 = wave_sum_f64(e16);
  if (lane == 0) {
    pack_bounds(a, row, PREC_F32, n32, e32);
    if (a.x16) pack_bounds(a, row, PREC_BF16, n16, e16);
  }
  if (lane < 4) {
    float v = 0.f;
    if (a.is_query) {
      v = (lane < 2 && real) ? 1.f : 0.f;
    } else if (real) {
      const double t = -0.5 * n32;
      const float hi = (float)t;
      v = lane == 0 ? hi : (lane == 1 ? (float)(t - (double)hi) : 0.f);
    }
    a.x32[row * a.ld32 + a.dp + lane] = v;
  }
  if (a.x16 && lane < 16) {
    uint16_t b = 0;
    if (real) {
      if (a.is_query) {
        b = lane < 3 ? f2bf(1.f) : 0;
      } else {
        const double t = -0.5 * n16;
        const uint16_t t1 = f2bf((float)t);
        const double r1 = t - (double)bf2f(t1);
        const uint16_t t2 = f2bf((float)r1);
        const double r2 = r1 - (double)bf2f(t2);
        const uint16_t t3 = f2bf((float)r2);
        b = lane == 0 ? t1 : lane == 1 ? t2 : lane == 2 ? t3 : 0;
      }
    }
    a.x16[row * a.ld16 + a.dp + lane] = b;
  }
}

// ------------------------------------------------------------------------------------------------
// FP8 (OCP E4M3: 1 sign, 4 exponent bits with bias 7, 3 mantissa bits; max 448, no infinities)
// ------------------------------------------------------------------------------------------------
// Round to nearest even, saturating at +-448 (inputs are scaled into range first).
__device__ __forceinline__ uint8_t f32_to_e4m3(float x) {
  const uint32_t u = __builtin_bit_cast(uint32_t, x);
  const uint8_t sgn = (uint8_t)((u >> 24) & 0x80u);
  const float ax = __builtin_fabsf(x);
  if (ax >= 448.0f) return sgn | 0x7Eu;
  if (ax < 0.015625f) {                       // below 2^-6: subnormals k * 2^-9, k = 0..8
    const int kq = (int)__builtin_rintf(ax * 512.0f);
    return sgn | (uint8_t)kq;                 // k = 8 is the smallest normal, 0x08
  }
  const uint32_t ua = __builtin_bit_cast(uint32_t, ax);
  uint32_t r = ua + 0x7FFFFu + ((ua >> 20) & 1u);   // keep 3 mantissa bits, ties to even
  r &= 0xFFF00000u;
  const int e = (int)((r >> 23) & 0xFF) - 127;
  const uint32_t mnt = (r >> 20) & 7u;
  if (e > 8 || (e == 8 && mnt == 7u)) return sgn | 0x7Eu;
  return sgn | (uint8_t)(((uint32_t)(e + 7) << 3) | mnt);
}
__device__ __forceinline__ double e4m3_to_f64(uint8_t b) {
  const int ex = (b >> 3) & 15, mnt = b & 7;
  const double v = ex == 0 ? (double)mnt * 0x1p-9 : (1.0 + mnt / 8.0) * __builtin_ldexp(1.0, ex - 7);
  return (b & 0x80) ? -v : v;
}
// smallest power-of-two exponent p with amax / 2^p <= 448 (0 for amax = 0), clamped to E8M0
__device__ __forceinline__ int fp8_scale_exp(double amax) {
  if (!(amax > 0.0)) return 0;
  const uint64_t ub = __builtin_bit_cast(uint64_t, amax / 448.0);   // ceil(log2) from the bits
  const int eb = (int)((ub >> 52) & 0x7FF) - 1023;
  int p = (ub & ((1ull << 52) - 1)) ? eb + 1 : eb;
  while (amax / __builtin_ldexp(1.0, p) > 448.0) ++p;
  while (p > -126 && amax / __builtin_ldexp(1.0, p - 1) <= 448.0) --p;
  return p < -126 ? -126 : (p > 126 ? 126 : p);
}

// One wave per row: wave max -> scale, quantise, FP64 norm of the dequantised row, aug unit.
__global__ __launch_bounds__(256) void pack_fp8_kernel(PackArgs a) {
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= a.rows_pad) return;
  const bool real = row < a.rows;
  auto val = [&](int j) -> double {
    if (!real || j >= a.d) return 0.0;
    return a.src_f64 ? ((const double*)a.src)[row * a.ld_src + j]
                     : (double)((const float*)a.src)[row * a.ld_src + j];
  };
  double amax = 0.0;
  for (int j = lane; j < a.d; j += 64) amax = __builtin_fmax(amax, __builtin_fabs((double)(float)val(j)));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) amax = __builtin_fmax(amax, __shfl_xor(amax, o, 64));
  const int pe = fp8_scale_exp(amax);
  const double sc = __builtin_ldexp(1.0, pe), inv = __builtin_ldexp(1.0, -pe);
  uint8_t* out = a.x8 + row * a.ld8;
  double nrm = 0.0, err = 0.0;
  for (int j = lane; j < a.dp8; j += 64) {
    const double v = val(j);
    const float f = (float)v;
    const uint8_t b = f32_to_e4m3((float)((double)f * inv));
    out[j] = b;
    const double dq = e4m3_to_f64(b) * sc;
    nrm += dq * dq;
    err += (dq - v) * (dq - v);
  }
  nrm = wave_sum_f64(nrm);
  err = wave_sum_f64(err);
  if (lane == 0) pack_bounds(a, row, PREC_FP8, nrm, err);
  if (lane < 4) {
    // aug unit, as 4 dwords: the BF16 terms (t1, t2), (t3, 0) of -||r~||^2 / 2 (corpus rows;
    // zeros for queries), then e + 127 in byte 8, then zeros
    uint32_t w = 0;
    if (lane < 2 && !a.is_query && real) {
      const double t = -0.5 * nrm;
      const uint16_t t1 = f2bf((float)t);
      const double r1 = t - (double)bf2f(t1);
      const uint16_t t2 = f2bf((float)r1);
      const uint16_t t3 = f2bf((float)(r1 - (double)bf2f(t2)));
      w = lane == 0 ? ((uint32_t)t1 | ((uint32_t)t2 << 16)) : (uint32_t)t3;
    }
    if (lane == 2) w = (uint32_t)(pe + 127);
    *(uint32_t*)(out + a.dp8 + 4 * lane) = w;
  }
}

// ------------------------------------------------------------------------------------------------
// scan kernel
// ------------------------------------------------------------------------------------------------
// One wave owns QB blocks of 32 queries (MFMA columns) and one corpus slice; it streams the
// slice in tiles of 64 rows (two 32-row MFMA blocks). MFMA 32x32 D layout: lane l holds column
// l%32 and rows 8*(v/4) + 4*(l/32) + v%4 for v = 0..15. K is permuted so that lane half h holds
// the contiguous dims [c*DC + h*DC/2, c*DC + (h+1)*DC/2) of chunk c, identically for A and B.
//
// After each tile, every (query, row) with s = -D < threshold is appended to the lane's private
// sub-pool (query, slice, half) -- no atomics. A sub-pool close to full triggers the slow path:
// the wave computes the exact K-th entry of the query's two sub-pools, keeps exactly K entries
// and tightens that query's threshold.
template <int QB>
struct ScanState {
  float phi[QB];        // D-space threshold: admit D > phi  (phi = -threshold, exact negation)
  int cnt[QB];          // this lane's sub-pool fill
  int q0;               // first query of this wave's group
  int sl;               // slice
  // rarely used values are recomputed where needed instead of occupying registers
  __device__ __forceinline__ long pbase(const ScanArgs& a, int qb) const {
    const int q = q0 + qb * 32 + (int)(threadIdx.x & 31);
    return (((long)q * a.slices_max + sl) * 2 + (int)((threadIdx.x & 63) >> 5)) * a.sub_cap;
  }
  __device__ __forceinline__ uint32_t thr(int qb) const { return f2ord(-phi[qb]); }
};

// Sub-pool entries are written by the scan as (bits of D) << 32 | row, D = q~.r~ - ||r~||^2/2 the
// raw accumulator: the hot path spends no instruction on the key transform. Readers order them by
// pool_key: (ord(s) << 32) | row with s = -D, the canonical (key, index) order of the carry.
__device__ __forceinline__ uint64_t pool_key(uint64_t e) {
  const float D = __builtin_bit_cast(float, (uint32_t)(e >> 32));
  return ((uint64_t)f2ord(-D + 0.0f) << 32) | (e & 0xffffffffull);
}

// Slow path: the lanes whose sub-pool fill exceeds sub_cap - room select, per query, the exact
// K-th entry of the query's two sub-pools (K + 1 when the self row may be among them: the merge
// drops it later), keep the entries up to it and tighten that query's threshold.
template <int QB>
__device__ __forceinline__ void scan_compact(const ScanArgs& a, ScanState<QB>& st, int g, int sl,
                                          uint32_t* lds, int room) {
  const int lane = lane_opaque();
  const int lim = a.sub_cap - room;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const uint64_t need = __builtin_amdgcn_ballot_w64(st.cnt[qb] > lim);
    uint32_t cols = (uint32_t)(need | (need >> 32));
    while (cols) {
      const int c = __ffs(cols) - 1;
      cols &= cols - 1;
      const int q = (g * QB + qb) * 32 + c;
      const int n0 = __builtin_amdgcn_readlane(st.cnt[qb], c);
      const int n1 = __builtin_amdgcn_readlane(st.cnt[qb], c + 32);
      const int K = a.kc[q] + (a.self_ids ? 1 : 0);
      uint64_t* p0 = a.sub + (((long)q * a.slices_max + sl) * 2 + 0) * a.sub_cap;
      uint64_t* p1 = p0 + a.sub_cap;
      const int N = n0 + n1;
      if (lane == 0) atomicAdd(&a.flags[1], 1);
      if (N <= K) continue;
      auto raw = [&](int i) -> uint64_t { return i < n0 ? p0[i] : p1[i - n0]; };
      auto get = [&](int i) -> uint64_t { return pool_key(raw(i)); };
      const uint64_t E = group_kth<true>(get, N, K, lds);
      // the band above the K-th entry is kept when it fits in half the pool; otherwise exactly the
      // K entries, and thr[q] is lowered to their K-th key so that no later threshold of this
      // search admits the rows dropped here
      uint32_t T = (uint32_t)(E >> 32);
      bool banded = false;
      if (a.band.on) {
        const uint32_t fT = band_thr(a.band, q, T);
        int cnt = 0;
        for (int base = 0; base < N; base += 64) {
          const int i = base + lane;
          bool keep = false;
          if (i < N) { const uint64_t e = get(i); keep = (uint32_t)(e >> 32) < fT || e <= E; }
          cnt += __popcll(__builtin_amdgcn_ballot_w64(keep));
        }
        if (cnt <= lim / 2) {
          T = fT;
          banded = true;
        } else if (lane == 0) {
          atomicMin(&a.thr_w[q], T);
        }
      }
      // in-place order-preserving compaction of the logical array [p0 | p1] (writes never
      // overtake reads: a chunk is read completely before any of its outputs is stored)
      int outpos = 0;
      for (int base = 0; base < N; base += 64) {
        const int i = base + lane;
        uint64_t e = 0;
        bool keep = false;
        if (i < N) {
          e = raw(i);
          const uint64_t pk = pool_key(e);
          keep = pk <= E || (banded && (uint32_t)(pk >> 32) < T);
        }
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        const int pos = outpos + __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32),
                                  __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
        __builtin_amdgcn_wave_barrier();
        if (keep) { if (pos < n0) p0[pos] = e; else p1[pos - n0] = e; }
        outpos += __popcll(km);
      }
      const int k0 = outpos < n0 ? outpos : n0;
      const int k1 = outpos - k0;
      if (lane == c) { st.cnt[qb] = k0; }
      if (lane == c + 32) { st.cnt[qb] = k1; }
      if ((lane & 31) == c) {
        if (T < st.thr(qb)) st.phi[qb] = -ord2f(T);
      }
      if (k0 > lim || k1 > lim) {
        // no room for another tile: stop admitting for this query and report the overflow
        // (results of this search are then rejected by the host)
        if ((lane & 31) == c) st.phi[qb] = __builtin_inff();
        if (lane == 0) atomicAdd(&a.flags[0], 1);
      }
    }
  }
}

// Filter of one 64-row tile. Survivors are rare in the large rounds (~0.1% of the pairs), so the
// common case must cost a few VALU ops per value and no exec-mask juggling: each lane reduces its
// values in groups of 4 consecutive rows (v_max3 + v_max), one wave-uniform ballot per query
// block skips the block, one per group skips the group, and only groups holding a survivor run
// the per-value append (an address, a compare and a carry per value: the entry is the raw
// accumulator and the row, and the self row of an all-kNN query is dropped by the merge).
template <int QB, bool FULL>
__device__ __forceinline__ void scan_epilogue(const ScanArgs& a, ScanState<QB>& st,
                                              const f32x16 (&acc)[QB][2], long t0, long c1) {
  const int h4 = ((threadIdx.x & 63) >> 5) * 4;
  const int lim = FULL ? 64 : (int)(c1 - t0);   // valid rows of this tile
  const uint32_t tb = (uint32_t)t0 + (uint32_t)h4;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    float Dv[2][16];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rel = cb * 32 + 8 * (v >> 2) + (v & 3) + h4;
        Dv[cb][v] = (FULL || rel < lim) ? acc[qb][cb][v] : -__builtin_inff();
      }
    float gm[8];
#pragma unroll
    for (int gi = 0; gi < 8; ++gi) {
      const int cb = gi >> 2, v0 = (gi & 3) * 4;
      gm[gi] = fmaxf(fmaxf(Dv[cb][v0], Dv[cb][v0 + 1]), fmaxf(Dv[cb][v0 + 2], Dv[cb][v0 + 3]));
    }
    const float qm = fmaxf(fmaxf(fmaxf(gm[0], gm[1]), fmaxf(gm[2], gm[3])),
                           fmaxf(fmaxf(gm[4], gm[5]), fmaxf(gm[6], gm[7])));
    const float phi = st.phi[qb];
    if (__builtin_amdgcn_ballot_w64(qm > phi) == 0) continue;
    uint64_t* pool = a.sub + st.pbase(a, qb);
    int cnt = st.cnt[qb];
#pragma unroll
    for (int gi = 0; gi < 8; ++gi) {
      if (__builtin_amdgcn_ballot_w64(gm[gi] > phi) == 0) continue;
      // some lane holds a survivor in this group of 4 rows: one compare per value, whose result
      // is the exec mask of its store (room for a whole stage is guaranteed before the stage)
      const int cb = gi >> 2, v0 = (gi & 3) * 4;
      const uint32_t rg = tb + (uint32_t)(cb * 32 + 8 * (gi & 3));
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float D = Dv[cb][v0 + t];
        if (D > phi) {
          pool[cnt] = ((uint64_t)__builtin_bit_cast(uint32_t, D) << 32) | (rg + (uint32_t)t);
          ++cnt;
        }
      }
    }
    st.cnt[qb] = cnt;
  }
}

// Dense round (every row a candidate, e.g. the first round, whose threshold is +inf): a lane's
// group of 4 consecutive rows goes to 4 consecutive entries of the query's (slice, half 0)
// sub-pool at the rows' slice-relative positions, as two 16-byte stores, with no filter and no
// counter. Rows past the slice end land inside the pool (a stage never exceeds sub_cap rows) and
// are not counted.
template <int QB>
__device__ __forceinline__ void scan_dense_store(const ScanArgs& a, const ScanState<QB>& st,
                                                 const f32x16 (&acc)[QB][2], long t0, long c0) {
  const int h4 = ((threadIdx.x & 63) >> 5) * 4;
  const uint32_t rb = (uint32_t)t0 + (uint32_t)h4;
  const long pos0 = t0 - c0 + h4;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = st.q0 + qb * 32 + (int)(threadIdx.x & 31);
    uint64_t* pool = a.sub + (((long)q * a.slices_max + st.sl) * 2) * a.sub_cap + pos0;
#pragma unroll
    for (int gi = 0; gi < 8; ++gi) {
      const int cb = gi >> 2, v0 = (gi & 3) * 4;
      const uint32_t rel = (uint32_t)(cb * 32 + 8 * (gi & 3));
      // plain float copies: a bit cast applied to a vector element expression reads element 0
      const float d0 = acc[qb][cb][v0], d1 = acc[qb][cb][v0 + 1];
      const float d2 = acc[qb][cb][v0 + 2], d3 = acc[qb][cb][v0 + 3];
      const u32x4 w0 = {rb + rel, __builtin_bit_cast(uint32_t, d0), rb + rel + 1, __builtin_bit_cast(uint32_t, d1)};
      const u32x4 w1 = {rb + rel + 2, __builtin_bit_cast(uint32_t, d2), rb + rel + 3, __builtin_bit_cast(uint32_t, d3)};
      *(u32x4*)(pool + rel) = w0;
      *(u32x4*)(pool + rel + 2) = w1;
    }
  }
}

// LDS stage: SR corpus rows x one chunk of DC dims + the augmentation. A BF16 row keeps only the
// half-0 augmentation (8 elements: the 3-term -||r~||^2/2 split); half-1 lanes read the same
// values and multiply them by the query's zero half-1 augmentation. The row stride is padded to an
// odd multiple of 4 dwords so the 16-byte MFMA operand reads of 8 (32 banks) or 16 (64 banks)
// consecutive rows hit distinct banks.
template <int PREC, int DC, int SR>
struct TileGeom {
  static constexpr int ESZ = PREC == PREC_F32 ? 4 : 2;
  static constexpr int AUG = PREC == PREC_F32 ? 4 : 8;
  static constexpr int VEC = 16 / ESZ;
  static constexpr int ROW_EL = DC + AUG;
  static constexpr int ROW_DW0 = ROW_EL * ESZ / 4;
  static constexpr int STRIDE_DW = ROW_DW0 + ((4 - ROW_DW0 % 8) + 8) % 8;
  static constexpr int SU = STRIDE_DW / 4;          // 16-byte units per LDS row
  static constexpr int UNITS = ROW_EL / VEC;        // units holding data
  static constexpr int MAIN_UNITS = DC / VEC;
  static constexpr int BLOCKS = SR * SU / 64;       // wave-wide DMA loads per stage
  static constexpr int WLOADS = (BLOCKS + 3) / 4;   // per wave
  static constexpr int BUF_DW = SR * STRIDE_DW;
  static_assert(STRIDE_DW % 8 == 4, "stride must be an odd multiple of 4 dwords");
  static_assert(SR * SU % 64 == 0, "stage must be whole wave loads");
};

// (workgroup row, slice) of this workgroup under the XCD-contiguous item order of
// ScanArgs::wg_rows; false for the grid's padding to a multiple of 8 (the caller returns before
// any barrier)
__device__ __forceinline__ bool scan_item(const ScanArgs& a, int& wgrow, int& sl) {
  const int items = a.wg_rows * a.slices, per_xcd = (items + 7) >> 3;
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  wgrow = item % a.wg_rows;
  sl = item / a.wg_rows;
  return item < items;
}

// One workgroup = 4 waves = 4 query groups (32*QB queries each) sharing one corpus slice. The
// slice is streamed in stages of SR rows x one DC chunk through two LDS buffers filled by DMA: the
// loads of stage s+1 are in flight while the waves run the MFMAs of stage s. A BF16 stage of 128
// rows is processed as two 64-row sub-tiles, halving barriers and waits per MFMA (single-chunk
// only: a sub-tile's accumulators do not outlive the stage).
template <int PREC, int QB, int DC, int SR>
__global__ __launch_bounds__(256, 2) void scan_kernel(ScanArgs a) {
  using G = TileGeom<PREC, DC, SR>;
  constexpr int NBUF = 2;
  constexpr int SUB = SR / 64;
  __shared__ __attribute__((aligned(16))) uint32_t tile[NBUF][G::BUF_DW];
  __shared__ __attribute__((aligned(16))) uint32_t kth_lds[4][256 + 32];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int col = lane & 31, h = lane >> 5;
  int wgrow, sl;   // workgroup row (4 query groups), slice
  if (!scan_item(a, wgrow, sl)) return;
  // every query of the workgroup is an inert slot of a padded (graph-captured) search
  if (a.n_dev && wgrow * 4 * QB * 32 >= *a.n_dev) return;   // workgroup-uniform, before any barrier
  const int g = wgrow * 4 + wv;   // query group of this wave
  const long c0 = a.row_begin + (long)sl * a.slice_rows;
  long c1 = c0 + a.slice_rows;
  if (c1 > a.row_end) c1 = a.row_end;

  ScanState<QB> st;
  st.q0 = g * QB * 32;
  st.sl = sl;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    st.phi[qb] = -ord2f(a.thr[q]);
    st.cnt[qb] = 0;
  }
  const int nch = a.dp / DC;
  const long ntiles = c0 < c1 ? (c1 - c0 + SR - 1) / SR : 0;
  const long nstage = ntiles * nch;

  typedef typename std::conditional<PREC == PREC_F32, float, uint16_t>::type T;
  const T* X = (const T*)a.X;
  const T* Q = (const T*)a.Qp;

  // ---- stage loader: global -> LDS by DMA (buffer_load_dwordx4 ... lds) ----
  // The buffer is a dense array of 16-byte units, SU per row. Load l of wave wv fills units
  // [(wv + 4l)*64, +64): lane -> (row, unit) -> global source; padding units re-read the row start.
  // Blocks past the end repeat the last block (same source, destination and bytes) so every wave
  // issues exactly WLOADS loads per stage.
  const uint32_t ldx32 = (uint32_t)a.ldx;   // < 2^16 elements: in-stage offsets fit 32 bits
  auto load_stage = [&](long s, int buf) {
    const long t0 = c0 + (s / nch) * SR;
    const int c = (int)(s % nch);
    // one buffer descriptor per stage (base = the stage's first row, range = its SR rows): no
    // access can leave the stage, and each load is a 32-bit lane offset instead of an address
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(X + t0 * a.ldx), (short)0, (int)((uint32_t)SR * ldx32 * (uint32_t)sizeof(T)), 0x00020000);
    // recompute the per-lane offsets at every stage instead of keeping WLOADS of them live across
    // the loop (the empty asm hides that `ln` equals the loop-invariant lane id)
    int ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int l = 0; l < G::WLOADS; ++l) {
      const int blk = wv + 4 * l < G::BLOCKS ? wv + 4 * l : G::BLOCKS - 1;
      const int U = blk * 64 + ln;
      const uint32_t r = (uint32_t)(U / G::SU), u = (uint32_t)(U % G::SU);
      const uint32_t off = r * ldx32 + (u < G::MAIN_UNITS ? (uint32_t)(c * DC) + u * G::VEC
                                      : (u < G::UNITS ? (uint32_t)a.dp + (u - G::MAIN_UNITS) * G::VEC : 0u));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (void __attribute__((address_space(3)))*)&tile[buf][blk * 256], 16,
                                               (int)(off * (uint32_t)sizeof(T)), 0, 0, 0);
    }
  };

  // ---- per-wave query operand (B) ----
  constexpr int NB = PREC == PREC_F32 ? DC / 2 : DC / 16;
  typedef typename std::conditional<PREC == PREC_F32, float, bf16x8>::type BT;
  BT bq[QB][NB];
  auto loadB = [&](int c) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const T* qp = Q + (long)((g * QB + qb) * 32 + col) * a.ldq + c * DC + h * (DC / 2);
      if constexpr (PREC == PREC_F32) {
#pragma unroll
        for (int j = 0; j < NB; j += 4) {
          const f32x4 t = *(const f32x4*)(qp + j);
          bq[qb][j] = t[0]; bq[qb][j + 1] = t[1]; bq[qb][j + 2] = t[2]; bq[qb][j + 3] = t[3];
        }
      } else {
#pragma unroll
        for (int j = 0; j < NB; ++j) bq[qb][j] = __builtin_bit_cast(bf16x8, *(const u32x4*)(qp + 8 * j));
      }
    }
  };
  // the query side of the augmentation is the same for every query: ones against the norm terms
  // (FP32: hi and lo in the two lane halves; BF16: three terms in half 0, zeros in half 1)
  BT baug;
  if constexpr (PREC == PREC_F32) {
    baug = 1.0f;
  } else {
    const uint32_t one = 0x3f80u, w01 = h ? 0u : (one | (one << 16)), w2 = h ? 0u : one;
    baug = __builtin_bit_cast(bf16x8, (u32x4){w01, w2, 0u, 0u});
  }
  if (nch == 1) loadB(0);

  if (nstage > 0) load_stage(0, 0);
  f32x16 acc[QB][2];
  for (long s = 0; s < nstage; ++s) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's DMA for stage s has landed
    __syncthreads();                 // ... every wave's; and nobody still reads the other buffer
    const int buf = (int)(s & 1);
    const int c = (int)(s % nch);
    const long t0s = c0 + (s / nch) * SR;
    if (s + 1 < nstage) load_stage(s + 1, buf ^ 1);
    if (nch > 1) loadB(c);
#pragma unroll
    for (int sub = 0; sub < SUB; ++sub) {
      const long t0 = t0s + 64 * sub;
      if (SUB > 1 && t0 >= c1) break;
      // one sub-tile's accumulators at a time (register budget of 2 waves/SIMD)
      __builtin_amdgcn_sched_barrier(0);
      if (c == 0) {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
#pragma unroll
          for (int cb = 0; cb < 2; ++cb) acc[qb][cb] = (f32x16){0};
      }
      const uint32_t* r0 = &tile[buf][(64 * sub + col) * G::STRIDE_DW];
      const uint32_t* r1 = &tile[buf][(64 * sub + 32 + col) * G::STRIDE_DW];
      const int hoff = h * (DC / 2) * G::ESZ / 4;  // dword offset of this lane half
      if constexpr (PREC == PREC_F32) {
#pragma unroll
        for (int j = 0; j < DC / 2; j += 4) {
          const f32x4 a0 = __builtin_bit_cast(f32x4, *(const u32x4*)(r0 + hoff + j));
          const f32x4 a1 = __builtin_bit_cast(f32x4, *(const u32x4*)(r1 + hoff + j));
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) {
              acc[qb][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[jj], bq[qb][j + jj], acc[qb][0], 0, 0, 0);
              acc[qb][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[jj], bq[qb][j + jj], acc[qb][1], 0, 0, 0);
            }
          }
        }
        if (c == nch - 1) {
          const float g0 = __builtin_bit_cast(float, r0[DC + h]);
          const float g1 = __builtin_bit_cast(float, r1[DC + h]);
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            acc[qb][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(g0, baug, acc[qb][0], 0, 0, 0);
            acc[qb][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(g1, baug, acc[qb][1], 0, 0, 0);
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < DC / 16; ++j) {
          const bf16x8 a0 = __builtin_bit_cast(bf16x8, *(const u32x4*)(r0 + hoff + 4 * j));
          const bf16x8 a1 = __builtin_bit_cast(bf16x8, *(const u32x4*)(r1 + hoff + 4 * j));
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            acc[qb][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bq[qb][j], acc[qb][0], 0, 0, 0);
            acc[qb][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bq[qb][j], acc[qb][1], 0, 0, 0);
          }
          // keep at most two steps of operands in flight (register budget of 2 waves/SIMD)
          if ((j & 1) == 1) __builtin_amdgcn_sched_barrier(0);
        }
        if (c == nch - 1) {
          // both lane halves read the half-0 augmentation; half 1 multiplies it by zeros
          const bf16x8 g0 = __builtin_bit_cast(bf16x8, *(const u32x4*)(r0 + DC / 2));
          const bf16x8 g1 = __builtin_bit_cast(bf16x8, *(const u32x4*)(r1 + DC / 2));
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            acc[qb][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g0, baug, acc[qb][0], 0, 0, 0);
            acc[qb][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g1, baug, acc[qb][1], 0, 0, 0);
          }
        }
      }
      if (c == nch - 1) {
        if (a.dense) scan_dense_store<QB>(a, st, acc, t0, c0);
        else if (t0 + 64 <= c1) scan_epilogue<QB, true>(a, st, acc, t0, c1);
        else scan_epilogue<QB, false>(a, st, acc, t0, c1);
        bool trig = false;
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) trig |= st.cnt[qb] > a.sub_cap - 32;
        if (__builtin_amdgcn_ballot_w64(trig)) scan_compact<QB>(a, st, g, sl, kth_lds[wv], 32);
      }
    }
  }
  // publish fills
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    const int fill = a.dense ? (h == 0 && c1 > c0 ? (int)(c1 - c0) : 0) : st.cnt[qb];
    a.sub_cnt[((long)q * a.slices_max + sl) * 2 + h] = fill;
  }
}

// ------------------------------------------------------------------------------------------------
// BF16 single-chunk scan (DP = DC in {32, 64, 128}, the headline configuration) and FP8 scan
// ------------------------------------------------------------------------------------------------
// Same work split and filter as scan_kernel; the stage layout is built so that the loop spends
// almost no VALU on addressing:
//   * an LDS row holds only the SU 16-byte units of main dims (no padding); unit u of row r sits
//     at position u ^ sw(r), sw(r) = (r / (16/SU)) % SU, so the 16-byte MFMA operand reads of 16
//     consecutive rows (8 with 32 banks) touch distinct banks;
//   * the augmentation units (16 bytes per row) go to a separate LDS array;
//   * a wave-wide DMA load (a buffer load into LDS through the stage's descriptor) fills 64
//     consecutive units = 64/SU whole rows, so with load L of wave wv being L = wv + 4l, a lane's
//     offset inside the stage is the same for every load and every stage (one register); loads
//     differ only by a scalar offset;
//   * operand reads for step j are at (A0 ^ 16j) + immediate, one XOR per step;
//   * the augmentation step (one MFMA per 32-row half with a zero C operand) is the first term of
//     every accumulator, so no accumulator is ever cleared; operands are read one step ahead of
//     their MFMAs.
// BF16 rows (SU = dp/8): one 32x32x16 MFMA per unit pair; lane half h takes units h*SU/2 + j.
// FP8 rows (OCP E4M3, SU = dp8/16): one scaled 32x32x64 MFMA per 4 units; lane half h takes units
// h*SU/2 + 2j and + 2j + 1 (32 bytes). The E8M0 row scales (2^e_r of the corpus row, 2^e_q of the
// query) are the MFMA's per-lane scale operands, so the products are those of the dequantised
// values r~ = 2^e_r * e4m3(r / 2^e_r). The augmentation stays a BF16 MFMA against the 3 BF16 terms
// of -||r~||^2 / 2 (half the issue cost of an FP8 one); its unit's byte 8 holds e_r + 127, which the
// BF16 step multiplies by zero.
template <int SU_>
struct SwzGeom {
  static constexpr int SU = SU_;                 // 16-byte units of main dims per row
  static constexpr int SR = 128;                 // rows per stage
  static constexpr int RPL = 64 / SU;            // rows per wave-wide DMA load
  static constexpr int MAIN_LOADS = SR / RPL;    // per stage
  static constexpr int PER_WAVE = MAIN_LOADS / 4;
  static constexpr int MAIN_BYTES = SR * SU * 16;
  static constexpr int AUG_BYTES = SR * 16;
  static_assert(SU == 4 || SU == 8 || SU == 16, "rows of 64, 128 or 256 bytes");
  static_assert(MAIN_LOADS % 4 == 0, "whole loads per wave");
};

typedef int i32x8 __attribute__((ext_vector_type(8)));

template <int QB, int SU, bool F8>
__global__ __launch_bounds__(256, QB >= 3 ? 1 : 2) void scan_swz_kernel(ScanArgs a) {
  using G = SwzGeom<SU>;
  constexpr int ESZ = F8 ? 1 : 2;                // bytes per element
  constexpr int J = F8 ? SU / 4 : SU / 2;        // MFMA k-steps per row block
  __shared__ __attribute__((aligned(1024))) uint8_t smem[2 * G::MAIN_BYTES + 2 * G::AUG_BYTES];
  __shared__ __attribute__((aligned(16))) uint32_t kth_lds[4][256 + 32];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar addressing
  const int col = lane & 31, h = lane >> 5;
  int wgrow, sl;   // workgroup row (4 query groups), slice
  if (!scan_item(a, wgrow, sl)) return;
  // every query of the workgroup is an inert slot of a padded (graph-captured) search
  if (a.n_dev && wgrow * 4 * QB * 32 >= *a.n_dev) return;   // workgroup-uniform, before any barrier
  const int g = wgrow * 4 + wv;
  // corpus rows are int32 indices: every row offset below fits 32 bits
  const int c0 = (int)(a.row_begin + (long)sl * a.slice_rows);
  int c1 = c0 + (int)a.slice_rows;
  if (c1 > (int)a.row_end) c1 = (int)a.row_end;

  ScanState<QB> st;
  st.q0 = g * QB * 32;
  st.sl = sl;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    st.phi[qb] = -ord2f(a.thr[q]);
    st.cnt[qb] = 0;
  }
  const int nst = c0 < c1 ? (c1 - c0 + G::SR - 1) / G::SR : 0;
  const uint8_t* X = (const uint8_t*)a.X;
  const uint32_t ldb = (uint32_t)a.ldx * ESZ;    // row stride in bytes

  // ---- DMA source offsets (bytes, relative to the first row of the load) ----
  const int swl = (4 * wv + (lane >> 4)) % SU;
  const uint32_t dma_off = (uint32_t)(lane / SU) * ldb + (uint32_t)((lane % SU) ^ swl) * 16u;
  const uint32_t aug_off = (uint32_t)lane * ldb + (uint32_t)a.dp * ESZ;
  typedef __attribute__((address_space(3))) uint8_t lds_u8;
  // All loads of a stage, issued right after the barrier (spreading them over the k-steps was
  // measured: no gain for BF16, 4% slower for FP8): buffer loads into LDS through one descriptor
  // per stage (base = the stage's first row, range = its SR rows, so no access can leave the
  // stage), a fixed per-lane offset and a scalar offset per load -- three instructions per load.
  // Load k < PER_WAVE carries the main units of rows wv*RPL + 4*RPL*k; waves 0 and 1 then load
  // the augmentation units of rows 64*wv .. 64*wv + 63. The empty asm keeps the LDS addresses
  // in SGPRs.
  const int kstride = (int)(4 * G::RPL) * (int)ldb;   // bytes between a wave's main loads
  const int voff = (int)(dma_off + (uint32_t)(wv * G::RPL) * ldb);
  const int aoff = (int)(aug_off + (uint32_t)(wv * 64) * ldb);
  auto issue = [&](int t0, int buf) {
    // word 3 = 0x00020000: 32-bit data format, the raw-buffer setting of CDNA descriptors
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(X + (size_t)t0 * ldb), (short)0, (int)(G::SR * ldb), 0x00020000);
    int soff = 0;
#pragma unroll
    for (int k = 0; k < G::PER_WAVE; ++k) {
      lds_u8* l = (lds_u8*)(smem + buf * G::MAIN_BYTES + wv * 1024 + k * 4096);
      asm volatile("" : "+s"(l));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (void __attribute__((address_space(3)))*)l, 16, voff, soff,
                                               0, 0);
      soff += kstride;
    }
    if (wv < G::SR / 64) {
      lds_u8* la = (lds_u8*)(smem + 2 * G::MAIN_BYTES + buf * G::AUG_BYTES + wv * 1024);
      asm volatile("" : "+s"(la));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (void __attribute__((address_space(3)))*)la, 16, aoff, 0, 0, 0);
    }
  };

  // ---- per-wave query operand (B), resident for the whole slice ----
  typedef typename std::conditional<F8, i32x8, bf16x8>::type OT;
  OT bq[QB][J];
  int sq[QB];                                    // FP8: E8M0 scale byte of the lane's query
  {
    const uint8_t* Q = (const uint8_t*)a.Qp;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const uint8_t* qrow = Q + (size_t)((g * QB + qb) * 32 + col) * a.ldq * ESZ;
      const uint8_t* qp = qrow + h * (SU / 2) * 16;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if constexpr (F8) {
          const u32x4 lo = *(const u32x4*)(qp + 32 * j), hi = *(const u32x4*)(qp + 32 * j + 16);
          bq[qb][j] = (i32x8){(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                              (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
        } else {
          bq[qb][j] = __builtin_bit_cast(bf16x8, *(const u32x4*)(qp + 16 * j));
        }
      }
      sq[qb] = F8 ? (int)(*(const uint32_t*)(qrow + a.dp + 8) & 0xFFu) : 0;
    }
  }
  bf16x8 baug;
  {
    const uint32_t one = 0x3f80u, w01 = h ? 0u : (one | (one << 16)), w2 = h ? 0u : one;
    baug = __builtin_bit_cast(bf16x8, (u32x4){w01, w2, 0u, 0u});
  }
  // ---- operand read offsets (bytes inside a buffer) ----
  const int swc = (col / (16 / SU)) % SU;
  const uint32_t A0 = (uint32_t)col * SU * 16u + (uint32_t)(((h * (SU / 2)) ^ swc) * 16);
  const uint32_t AG = (uint32_t)col * 16u;

  if (nst > 0) issue(c0, 0);
  f32x16 acc[QB][2];
  for (int s = 0; s < nst; ++s) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's DMA for stage s has landed
    __syncthreads();                 // ... every wave's; and nobody still reads the other buffer
    const int buf = s & 1;
    const int t0s = c0 + s * G::SR;
    if (s + 1 < nst) issue(t0s + G::SR, buf ^ 1);
    const uint8_t* mb = smem + buf * G::MAIN_BYTES;
    const uint8_t* ab = smem + 2 * G::MAIN_BYTES + buf * G::AUG_BYTES;
#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      const int t0 = t0s + 64 * sub;
      if (t0 >= c1) break;
      __builtin_amdgcn_sched_barrier(0);
      const uint8_t* m0 = mb + sub * 64 * SU * 16;
      const uint8_t* m1 = m0 + 32 * SU * 16;
      auto rd = [&](const uint8_t* p, uint32_t off) -> u32x4 { return *(const u32x4*)(p + off); };
      // operand of step j: one unit (BF16) or two (FP8; the second at offset ^ 16)
      auto rdop = [&](const uint8_t* p, uint32_t off) -> OT {
        if constexpr (F8) {
          const u32x4 lo = rd(p, off), hi = rd(p, off ^ 16u);
          return (i32x8){(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                         (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
        } else {
          return __builtin_bit_cast(bf16x8, rd(p, off));
        }
      };
      constexpr uint32_t STEP = F8 ? 32u : 16u;
      const u32x4 g0 = rd(ab + sub * 64 * 16, AG), g1 = rd(ab + sub * 64 * 16 + 32 * 16, AG);
      const int s0 = (int)(g0[2] & 0xFFu), s1 = (int)(g1[2] & 0xFFu);
      OT y0 = rdop(m0, A0), y1 = rdop(m1, A0);
      // the norm step does not depend on the queries: one MFMA per 32-row half into query block
      // 0's accumulator, which every query block's first main step takes as its C operand
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, g0), baug,
                                                          (f32x16){0}, 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, g1), baug,
                                                          (f32x16){0}, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const OT x0 = y0, x1 = y1;
        if (j + 1 < J) {
          const uint32_t o = A0 ^ (STEP * (uint32_t)(j + 1));
          y0 = rdop(m0, o); y1 = rdop(m1, o);
        }
        // descending: at j = 0 the blocks qb > 0 read block 0's norm result before block 0 adds to it
#pragma unroll
        for (int qb = QB - 1; qb >= 0; --qb) {
          const f32x16 c0 = j == 0 ? acc[0][0] : acc[qb][0];
          const f32x16 c1 = j == 0 ? acc[0][1] : acc[qb][1];
          if constexpr (F8) {
            // formats 0, 0 = E4M3 x E4M3; scale bytes in bits 0..7 (op_sel 0)
            acc[qb][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x0, bq[qb][j], c0, 0, 0,
                                                                         0, s0, 0, sq[qb]);
            acc[qb][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x1, bq[qb][j], c1, 0, 0,
                                                                         0, s1, 0, sq[qb]);
          } else {
            acc[qb][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0, bq[qb][j], c0, 0, 0, 0);
            acc[qb][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1, bq[qb][j], c1, 0, 0, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (a.dense) scan_dense_store<QB>(a, st, acc, t0, c0);
      else if (t0 + 64 <= c1) scan_epilogue<QB, true>(a, st, acc, t0, c1);
      else scan_epilogue<QB, false>(a, st, acc, t0, c1);
    }
    // room for the next stage: at most 64 appends per lane and query block
    bool trig = false;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) trig |= st.cnt[qb] > a.sub_cap - 64;
    if (__builtin_amdgcn_ballot_w64(trig)) scan_compact<QB>(a, st, g, sl, kth_lds[wv], 64);
  }
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    const int fill = a.dense ? (h == 0 && c1 > c0 ? (int)(c1 - c0) : 0) : st.cnt[qb];
    a.sub_cnt[((long)q * a.slices_max + sl) * 2 + h] = fill;
  }
}

// ------------------------------------------------------------------------------------------------
// Pipelined BF16 / FP8 threshold scan
// ------------------------------------------------------------------------------------------------
// The stage layout, DMA, query operands and filter rule of scan_swz_kernel, with each 128-row stage
// cut into four tiles of 32 rows (one row block x the wave's QB query blocks) that run as a
// software pipeline, so that LDS latency and the filter hide under MFMA issue:
//   * a tile's operand registers are refilled with the next tile's units right after their last
//     MFMA: a whole tile of MFMAs separates a read from its use (one k-step in scan_swz_kernel,
//     shorter than the LDS latency under load);
//   * the next tile's norm step is issued in the middle of the current chain, so its result is
//     ready when that tile starts; it is the C operand of every query block's first step;
//   * a tile's filter runs during the next tile's chain: its max tree in parts between the k-steps'
//     MFMAs (no branches), its appends after the chain. The last tile of a stage is filtered during
//     the first tile of the next stage (the room check between stages still covers four tiles).
// A scheduling barrier after each k-step keeps this order (the compiler otherwise sinks the reads
// next to their use).
// Rows past the slice end (a partial last stage) are computed, never appended: the max tree does
// not mask them (a false hit only enters the append path), the append checks each row.
// Dense rounds (no threshold yet) run scan_swz_kernel.
template <int QB>
struct TileHits {
  float gm[QB][4];    // per query block: the lane's maximum over each group of 4 consecutive rows
  uint64_t hit[QB];   // ballot over the wave: some lane of the query block holds a survivor
};

// Part `part` of NPARTS of a tile's max tree: an even share of its QB x 4 group maxima; the last
// part also reduces them and takes the ballots. The parts are spread over the k-steps of the next
// tile's MFMA chain (the loop is unrolled: the selection folds away).
template <int QB, int NPARTS>
__device__ __forceinline__ void tile_max_part(const ScanState<QB>& st, const f32x16 (&acc)[QB], TileHits<QB>& th,
                                              int part) {
  constexpr int NG = QB * 4;
#pragma unroll
  for (int x = 0; x < NG; ++x) {
    if (x * NPARTS / NG != part) continue;
    const int qb = x / 4, gi = x % 4;
    th.gm[qb][gi] = fmaxf(fmaxf(acc[qb][4 * gi], acc[qb][4 * gi + 1]),
                          fmaxf(acc[qb][4 * gi + 2], acc[qb][4 * gi + 3]));
  }
  if (part != NPARTS - 1) return;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const float qm = fmaxf(fmaxf(th.gm[qb][0], th.gm[qb][1]), fmaxf(th.gm[qb][2], th.gm[qb][3]));
    th.hit[qb] = __builtin_amdgcn_ballot_w64(qm > st.phi[qb]);
  }
}

// Appends of one tile: lane value v is row t0 + 8*(v/4) + (v%4) + 4*half; rows at or past `lim`
// (relative to t0) lie beyond the slice and are skipped.
template <int QB, bool FULL>
__device__ __forceinline__ void tile_append(const ScanArgs& a, ScanState<QB>& st, const f32x16 (&acc)[QB],
                                            const TileHits<QB>& th, uint32_t t0, int lim) {
  const int h4 = (lane_opaque() >> 5) * 4;
  const uint32_t tb = t0 + (uint32_t)h4;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    if (th.hit[qb] == 0) continue;
    const float phi = st.phi[qb];
    uint64_t* pool = a.sub + st.pbase(a, qb);
    int cnt = st.cnt[qb];
#pragma unroll
    for (int gi = 0; gi < 4; ++gi) {
      if (__builtin_amdgcn_ballot_w64(th.gm[qb][gi] > phi) == 0) continue;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float D = acc[qb][4 * gi + t];
        if ((FULL || 8 * gi + t + h4 < lim) && D > phi) {
          pool[cnt] = ((uint64_t)__builtin_bit_cast(uint32_t, D) << 32) | (tb + (uint32_t)(8 * gi + t));
          ++cnt;
        }
      }
    }
    st.cnt[qb] = cnt;
  }
}

template <int QB>
__device__ __forceinline__ void tile_filter(const ScanArgs& a, ScanState<QB>& st, const f32x16 (&acc)[QB],
                                            const TileHits<QB>& th, uint32_t t0, int c1) {
  const int lim = c1 - (int)t0;
  if (lim >= 32) tile_append<QB, true>(a, st, acc, th, t0, lim);
  else if (lim > 0) tile_append<QB, false>(a, st, acc, th, t0, lim);
}

// Two resident workgroups per CU (two waves per SIMD). Four query blocks of 128-dim rows with one
// workgroup per CU (one wave per SIMD, half the LDS traffic per MFMA) measured 3.34 ms per
// headline search against 2.49 ms for two blocks: one wave cannot cover the stage barrier, the
// DMA wait and the appends.
template <int QB, int SU, bool F8>
__global__ __launch_bounds__(256, 2) void scan_pipe_kernel(ScanArgs a) {
  using G = SwzGeom<SU>;
  constexpr int ESZ = F8 ? 1 : 2;                // bytes per element
  constexpr int J = F8 ? SU / 4 : SU / 2;        // MFMA k-steps per tile
  constexpr int TILES = G::SR / 32;
  constexpr uint32_t TILE_MAIN = 32 * SU * 16;   // LDS bytes of a tile's main units
  __shared__ __attribute__((aligned(1024))) uint8_t smem[2 * G::MAIN_BYTES + 2 * G::AUG_BYTES];
  __shared__ __attribute__((aligned(16))) uint32_t kth_lds[4][256 + 32];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar addressing
  const int col = lane & 31, h = lane >> 5;
  int wgrow, sl;   // workgroup row (4 query groups), slice
  if (!scan_item(a, wgrow, sl)) return;
  if (a.n_dev && wgrow * 4 * QB * 32 >= *a.n_dev) return;   // workgroup-uniform, before any barrier
  const int g = wgrow * 4 + wv;
  const int c0 = (int)(a.row_begin + (long)sl * a.slice_rows);
  int c1 = c0 + (int)a.slice_rows;
  if (c1 > (int)a.row_end) c1 = (int)a.row_end;

  ScanState<QB> st;
  st.q0 = g * QB * 32;
  st.sl = sl;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    st.phi[qb] = -ord2f(a.thr[q]);
    st.cnt[qb] = 0;
  }
  const int nst = c0 < c1 ? (c1 - c0 + G::SR - 1) / G::SR : 0;
  const uint8_t* X = (const uint8_t*)a.X;
  const uint32_t ldb = (uint32_t)a.ldx * ESZ;

  // corpus DMA: the loads, descriptor and offsets of scan_swz_kernel
  const int swl = (4 * wv + (lane >> 4)) % SU;
  const uint32_t dma_off = (uint32_t)(lane / SU) * ldb + (uint32_t)((lane % SU) ^ swl) * 16u;
  const uint32_t aug_off = (uint32_t)lane * ldb + (uint32_t)a.dp * ESZ;
  typedef __attribute__((address_space(3))) uint8_t lds_u8;
  const int kstride = (int)(4 * G::RPL) * (int)ldb;
  const int voff = (int)(dma_off + (uint32_t)(wv * G::RPL) * ldb);
  const int aoff = (int)(aug_off + (uint32_t)(wv * 64) * ldb);
  auto issue = [&](int t0, int buf) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(X + (size_t)t0 * ldb), (short)0, (int)(G::SR * ldb), 0x00020000);
    int soff = 0;
#pragma unroll
    for (int k = 0; k < G::PER_WAVE; ++k) {
      lds_u8* l = (lds_u8*)(smem + buf * G::MAIN_BYTES + wv * 1024 + k * 4096);
      asm volatile("" : "+s"(l));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (void __attribute__((address_space(3)))*)l, 16, voff, soff,
                                               0, 0);
      soff += kstride;
    }
    if (wv < G::SR / 64) {
      lds_u8* la = (lds_u8*)(smem + 2 * G::MAIN_BYTES + buf * G::AUG_BYTES + wv * 1024);
      asm volatile("" : "+s"(la));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (void __attribute__((address_space(3)))*)la, 16, aoff, 0, 0, 0);
    }
  };

  // per-wave query operand (B), resident for the whole slice
  typedef typename std::conditional<F8, i32x8, bf16x8>::type OT;
  OT bq[QB][J];
  int sq[QB];
  {
    const uint8_t* Q = (const uint8_t*)a.Qp;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const uint8_t* qrow = Q + (size_t)((g * QB + qb) * 32 + col) * a.ldq * ESZ;
      const uint8_t* qp = qrow + h * (SU / 2) * 16;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if constexpr (F8) {
          const u32x4 lo = *(const u32x4*)(qp + 32 * j), hi = *(const u32x4*)(qp + 32 * j + 16);
          bq[qb][j] = (i32x8){(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                              (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
        } else {
          bq[qb][j] = __builtin_bit_cast(bf16x8, *(const u32x4*)(qp + 16 * j));
        }
      }
      sq[qb] = F8 ? (int)(*(const uint32_t*)(qrow + a.dp + 8) & 0xFFu) : 0;
    }
  }
  bf16x8 baug;
  {
    const uint32_t one = 0x3f80u, w01 = h ? 0u : (one | (one << 16)), w2 = h ? 0u : one;
    baug = __builtin_bit_cast(bf16x8, (u32x4){w01, w2, 0u, 0u});
  }
  // operand read offsets: a tile's rows start at a multiple of 32, so the swizzle of its row col
  // depends on col only, as in scan_swz_kernel
  const int swc = (col / (16 / SU)) % SU;
  const uint32_t A0 = (uint32_t)col * SU * 16u + (uint32_t)(((h * (SU / 2)) ^ swc) * 16);
  const uint32_t AG = (uint32_t)col * 16u;
  constexpr uint32_t STEP = F8 ? 32u : 16u;
  auto rd = [&](const uint8_t* p, uint32_t off) -> u32x4 { return *(const u32x4*)(p + off); };
  auto rdop = [&](const uint8_t* p, uint32_t off) -> OT {
    if constexpr (F8) {
      const u32x4 lo = rd(p, off), hi = rd(p, off ^ 16u);
      return (i32x8){(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                     (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
    } else {
      return __builtin_bit_cast(bf16x8, rd(p, off));
    }
  };
  auto mma = [&](const OT& x, int qb, int j, const f32x16& c, int sx) -> f32x16 {
    if constexpr (F8) {
      // formats 0, 0 = E4M3 x E4M3; scale bytes in bits 0..7 (op_sel 0)
      return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x, bq[qb][j], c, 0, 0, 0, sx, 0, sq[qb]);
    } else {
      return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, bq[qb][j], c, 0, 0, 0);
    }
  };
  auto norm_step = [&](const u32x4& gu) -> f32x16 {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, gu), baug, (f32x16){0}, 0, 0, 0);
  };

  // OVL: a tile is filtered during the next tile's chain (two accumulator sets). Three query blocks
  // of 16-unit rows, and four blocks of 8-unit rows, leave no registers for a second set: each tile
  // is filtered right after its own chain, while the next tile's norm step runs (and the SIMD's
  // other wave keeps the MFMAs busy). Four blocks of 16-unit rows do not fit in 256 registers.
  constexpr bool OVL = QB <= 2 || (QB == 3 && SU <= 8);
  OT xa[J];                 // operands of the tile being multiplied
  // without OVL: operands read PD k-steps ahead of their MFMAs (crossing into the next tile)
  constexpr int PD = OVL ? 0 : (J >= 2 ? 2 : 1);
  u32x4 xg;                 // its augmentation unit (FP8: byte 8 = the row's scale exponent)
  f32x16 nrm;               // norm step of the tile (read by its first k-step only)
  f32x16 acc[OVL ? 2 : 1][QB];   // accumulators (ping-pong by tile parity when OVL)
  TileHits<QB> th;
  // the tile before the first one: nothing to filter
  if constexpr (OVL) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) acc[1][qb] = (f32x16){0};
  }
  uint32_t prev_t0 = (uint32_t)c0;
  int prev_c1 = c0;         // lim = prev_c1 - prev_t0 <= 0: no appends

  // the previous tile's max tree runs in NPARTS parts at k-steps 1 .. J-1 (step 0 when J = 1): its
  // last MFMAs have finished by then
  constexpr int NPARTS = J > 1 ? J - 1 : 1;
  if (nst > 0) issue(c0, 0);
  for (int s = 0; s < nst; ++s) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's DMA for stage s has landed
    __syncthreads();                 // ... every wave's; and nobody still reads the other buffer
    const int buf = s & 1;
    const int t0s = c0 + s * G::SR;
    if (s + 1 < nst) issue(t0s + G::SR, buf ^ 1);
    const uint8_t* mb = smem + buf * G::MAIN_BYTES;
    const uint8_t* ab = smem + 2 * G::MAIN_BYTES + buf * G::AUG_BYTES;
    xg = rd(ab, AG);
#pragma unroll
    for (int j = 0; j < (OVL ? J : PD); ++j) xa[j] = rdop(mb, A0 ^ (STEP * (uint32_t)j));
    // the norm step; without OVL it goes straight into block 0's accumulator (no register of its own)
    if constexpr (OVL) nrm = norm_step(xg);
    else acc[0][0] = norm_step(xg);
    int sx = F8 ? (int)(xg[2] & 0xFFu) : 0;
#pragma unroll
    for (int rb = 0; rb < TILES; ++rb) {
      const int cur = OVL ? rb & 1 : 0, pv = OVL ? cur ^ 1 : 0;
      const bool more = rb + 1 < TILES;
      const uint8_t* mn = mb + (rb + 1) * TILE_MAIN;   // the next tile's main units (if more)
      int sx_next = 0;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if (more && j == 0) xg = rd(ab + (rb + 1) * 32 * 16, AG);
#pragma unroll
        // descending: without OVL, at j = 0 the blocks qb > 0 read block 0's norm before block 0
        // adds to it
        for (int qb = QB - 1; qb >= 0; --qb)
          acc[cur][qb] = mma(xa[j], qb, j, j == 0 ? (OVL ? nrm : acc[cur][0]) : acc[cur][qb], sx);
        if constexpr (OVL) {
          if (more) xa[j] = rdop(mn, A0 ^ (STEP * (uint32_t)j));
        } else {
          // PD k-steps ahead (PD x QB MFMAs cover the read): PD + 1 operand sets live instead of J
          if (j + PD < J) xa[j + PD] = rdop(mb + rb * TILE_MAIN, A0 ^ (STEP * (uint32_t)(j + PD)));
          else if (more) xa[j + PD - J] = rdop(mn, A0 ^ (STEP * (uint32_t)(j + PD - J)));
        }
        if (more && j == J / 2) {
          if constexpr (OVL) nrm = norm_step(xg);
          if constexpr (F8) sx_next = (int)(xg[2] & 0xFFu);
        }
        if (OVL && (J == 1 || j > 0)) tile_max_part<QB, NPARTS>(st, acc[pv], th, J == 1 ? 0 : j - 1);
        // keep this order: reads a whole tile ahead of their use, the max tree between MFMAs
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (OVL) {
        // the previous tile's appends (the last tile of the previous stage when rb = 0)
        if (rb > 0) tile_filter<QB>(a, st, acc[pv], th, (uint32_t)(t0s + 32 * (rb - 1)), c1);
        else tile_filter<QB>(a, st, acc[pv], th, prev_t0, prev_c1);
      } else {
#pragma unroll
        for (int p = 0; p < NPARTS; ++p) tile_max_part<QB, NPARTS>(st, acc[cur], th, p);
        tile_filter<QB>(a, st, acc[cur], th, (uint32_t)(t0s + 32 * rb), c1);
        if (more) acc[0][0] = norm_step(xg);   // the next tile's norm, once this tile is filtered
      }
      if (more) sx = sx_next;
    }
    prev_t0 = (uint32_t)(t0s + 32 * (TILES - 1));
    prev_c1 = c1;
    // room until the next check: this stage's last tile and the next stage's first three, at most
    // 4 x 16 appends per lane and query block
    bool trig = false;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) trig |= st.cnt[qb] > a.sub_cap - 64;
    if (__builtin_amdgcn_ballot_w64(trig)) scan_compact<QB>(a, st, g, sl, kth_lds[wv], 64);
  }
  if constexpr (OVL) {   // the last tile
#pragma unroll
    for (int p = 0; p < NPARTS; ++p) tile_max_part<QB, NPARTS>(st, acc[1], th, p);
    tile_filter<QB>(a, st, acc[1], th, prev_t0, prev_c1);
  }
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int q = (g * QB + qb) * 32 + col;
    a.sub_cnt[((long)q * a.slices_max + sl) * 2 + h] = st.cnt[qb];
  }
}

// ------------------------------------------------------------------------------------------------
// merge kernel: one wave per query (4 queries per 256-thread block, no block barriers)
// ------------------------------------------------------------------------------------------------
constexpr int MERGE_MAX_SRC = 1 + 2 * 128;
constexpr int MW_ENT = 1024;   // entries gathered to LDS per wave (8 KB)

// Rank of the entry that sets a merge's threshold. The last merge keeps the K' smallest entries;
// before it, with the certificate band on, the merge uses the predicted rank of the K'-th entry in
// the prefix seen so far (predicted_rank). Its threshold is still the key of real rows, so every
// threshold stays a valid filter; when a prefix holds more of a query's K' nearest rows than
// predicted (a corpus ordered by that query's neighbourhood), the certificate fails for the query
// and it is searched again without predicted ranks. On random order this cuts the candidates of
// the early rounds about fourfold (1M x 128 mixture, k = 100: 1813 -> ~450 per query).
__device__ __forceinline__ int merge_rank(const MergeArgs& a, int K) {
  return a.band.on ? predicted_rank(K, a.seen) : K;
}

// New threshold of a merge whose K'-th entry is E: its key, or with the certificate band f(key),
// never above the query's current threshold (which an in-scan truncation may have lowered).
__device__ __forceinline__ uint32_t merge_theta(const MergeArgs& a, int q, uint64_t E) {
  const uint32_t tE = (uint32_t)(E >> 32);
  if (!a.band.on) return tE;
  const uint32_t f = band_thr(a.band, q, tE);
  const uint32_t cur = a.thr[q];
  return f < cur ? f : (cur > tE ? cur : tE);
}

__device__ __forceinline__ void ce_u64(uint64_t& x, uint64_t& y, bool up) {
  const bool sw = (x > y) == up;
  const uint64_t lo = sw ? y : x;
  y = sw ? x : y;
  x = lo;
}

// Ascending bitonic sort of P (a power of two, 64 <= P <= 1024) u64 entries in LDS by one wave.
// Stage (size, stride) compares entries pos and pos + stride (pos with bit `stride` clear) and
// orders them ascending when pos & size == 0, descending otherwise. Stages of stride >= 8 run
// through LDS; every stage of stride < 8 runs in registers on 8-entry blocks (block b = entries
// 8b .. 8b+7, read and written as 16-byte units) -- one block pass per level instead of three
// LDS stages, so a 1024-entry sort makes 28 LDS stages and 8 block passes instead of 55.
__device__ void wave_bitonic_u64_r8(uint64_t* v, int P) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63;
  const int nblk = P >> 3;
  auto load8 = [&](int b, uint64_t (&r)[8]) {
    const u64x2* p = (const u64x2*)(v + 8 * b);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u64x2 t = p[k];
      r[2 * k] = t[0];
      r[2 * k + 1] = t[1];
    }
  };
  auto store8 = [&](int b, const uint64_t (&r)[8]) {
    u64x2* p = (u64x2*)(v + 8 * b);
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = (u64x2){r[2 * k], r[2 * k + 1]};
  };
  // strides 4, 2, 1 of one level whose direction is uniform over the block
  auto tail3 = [&](uint64_t (&r)[8], bool up) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ce_u64(r[i], r[i + 4], up);
#pragma unroll
    for (int i : {0, 1, 4, 5}) ce_u64(r[i], r[i + 2], up);
#pragma unroll
    for (int i : {0, 2, 4, 6}) ce_u64(r[i], r[i + 1], up);
  };
  wave_sync();
  // levels 2, 4 and 8 lie inside the blocks
  for (int b = lane; b < nblk; b += 64) {
    uint64_t r[8];
    load8(b, r);
    ce_u64(r[0], r[1], true); ce_u64(r[2], r[3], false); ce_u64(r[4], r[5], true); ce_u64(r[6], r[7], false);
    ce_u64(r[0], r[2], true); ce_u64(r[1], r[3], true); ce_u64(r[4], r[6], false); ce_u64(r[5], r[7], false);
    ce_u64(r[0], r[1], true); ce_u64(r[2], r[3], true); ce_u64(r[4], r[5], false); ce_u64(r[6], r[7], false);
    tail3(r, (b & 1) == 0);
    store8(b, r);
  }
  for (int size = 16; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride >= 8; stride >>= 1) {
      wave_sync();
      for (int t = lane; t < P / 2; t += 64) {
        const int pos = 2 * t - (t & (stride - 1));
        const bool up = ((pos & size) == 0);
        const uint64_t x = v[pos], y = v[pos + stride];
        if ((x > y) == up) { v[pos] = y; v[pos + stride] = x; }
      }
    }
    wave_sync();
    for (int b = lane; b < nblk; b += 64) {
      uint64_t r[8];
      load8(b, r);
      tail3(r, ((8 * b) & size) == 0);
      store8(b, r);
    }
  }
  wave_sync();
}

__global__ __launch_bounds__(256) void merge_kernel(MergeArgs a) {
  __shared__ __attribute__((aligned(16))) uint64_t ent_all[4][MW_ENT];
  __shared__ int off_all[4][MERGE_MAX_SRC + 1];
  __shared__ __attribute__((aligned(16))) uint32_t kth_all[4][256 + 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + wv;
  if (q >= (a.n_dev ? *a.n_dev : a.n)) return;   // wave-uniform; only wave-level sync below
  uint64_t* ent = ent_all[wv];
  int* off = off_all[wv];
  const int nsrc = 1 + 2 * a.slices;
  const long subq = (long)q * a.slices_max * 2;
  // source fills -> exclusive offsets: lane l owns sources [l*P, l*P + P)
  {
    const int P = (nsrc + 63) / 64;
    int c[5] = {0, 0, 0, 0, 0};
    int tot = 0;
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int sidx = lane * P + u;
      if (u < P && sidx < nsrc) c[u] = sidx == 0 ? a.carry_in_cnt[q] : a.sub_cnt[subq + sidx - 1];
      tot += c[u];
    }
    const int excl = (int)wave_incl_scan_u32((uint32_t)tot, lane) - tot;
    int acc = excl;
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int sidx = lane * P + u;
      if (u < P && sidx < nsrc) off[sidx] = acc;
      acc += c[u];
    }
    if (lane == 63) off[nsrc] = acc;
    wave_sync();
  }
  const int N = off[nsrc];
  const int K = merge_rank(a, a.kc[q]);
  const uint64_t* cin = a.carry_in + (long)q * a.carry_cap;
  const uint64_t* sub = a.sub + subq * a.sub_cap;
  // the self row of an all-kNN query (appended by the scan like any row) maps to the largest key
  // and is never kept: the carry never holds it
  const uint32_t self = a.self_ids ? (uint32_t)a.self_ids[q] : 0xffffffffu;
  auto gget = [&](int i) -> uint64_t {
    int lo = 0, hi = nsrc - 1;     // last source whose offset is <= i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const int j = i - off[lo];
    if (lo == 0) return cin[j];
    const uint64_t e = sub[(long)(lo - 1) * a.sub_cap + j];
    return (uint32_t)e == self ? ~0ull : pool_key(e);
  };
  uint64_t* cout = a.carry_out + (long)q * a.carry_cap;
  uint32_t theta = 0;
  bool banded = false;   // keep the band below theta as well as the entries up to E
  // the entries up to E first, then (banded) the band beyond E below theta: the rerank takes the
  // carry's first K' entries as the shortlist proper
  auto compact = [&](auto get, uint64_t E) {
    int outpos = 0;
    for (int pass = 0; pass < (banded ? 2 : 1); ++pass) {
      for (int base = 0; base < N; base += 64) {
        const int i = base + lane;
        uint64_t e = 0;
        bool keep = false;
        if (i < N) {
          e = get(i);
          keep = e != ~0ull && (pass == 0 ? e <= E : (e > E && (uint32_t)(e >> 32) < theta));
        }
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        const int pos = outpos + __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
        if (keep && pos < a.carry_cap) cout[pos] = e;
        outpos += __popcll(km);
      }
    }
    return outpos;
  };
  if (N <= K) {
    const int kept = compact(gget, ~0ull);   // everything but a dropped self row
    if (lane == 0) a.carry_out_cnt[q] = kept;
    return;
  }
  uint64_t E;
  int kept;
  if (N <= MW_ENT) {
    // gather, branch-free: every slot computes a valid address (slots past N read the first
    // carry entry), so the four loads per lane are in flight together; conversion afterwards
    for (int base = 0; base < N; base += 256) {
      uint64_t v[4];
      int src[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + u * 64 + lane;
        const int ii = i < N ? i : 0;
        int lo = 0, hi = nsrc - 1;     // last source whose offset is <= ii
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (off[mid] <= ii) lo = mid; else hi = mid - 1;
        }
        const int jj = ii - off[lo];
        src[u] = i < N ? lo : -1;
        v[u] = *(lo == 0 ? cin + jj : sub + (long)(lo - 1) * a.sub_cap + jj);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t e = v[u];
        v[u] = src[u] <= 0 ? e : ((uint32_t)e == self ? ~0ull : pool_key(e));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + u * 64 + lane;
        if (i < N) ent[i] = v[u];
      }
    }
    wave_sync();
    if (a.bitonic) {
      int P = 64;
      while (P < N) P <<= 1;
      for (int i = N + lane; i < P; i += 64) ent[i] = ~0ull;
      wave_bitonic_u64_r8(ent, P);   // P >= 64
      E = ent[K - 1];
      theta = merge_theta(a, q, E);
      // sorted: the band below theta is a prefix; it extends the first K entries, up to the carry
      kept = K;
      if (a.band.on) {
        int c = 0;
        for (int base = 0; base < N; base += 64)
          c += __popcll(__builtin_amdgcn_ballot_w64(base + lane < N && ent[base + lane] != ~0ull &&
                                                    (uint32_t)(ent[base + lane] >> 32) < theta));
        kept = c > K ? c : K;
        if (kept > a.carry_cap) {
          kept = a.carry_cap;
          const uint32_t tl = (uint32_t)(ent[kept - 1] >> 32);
          theta = theta < tl ? theta : tl;
        }
      }
      for (int i = lane; i < kept; i += 64) cout[i] = ent[i];
    } else {
      auto lget = [&](int i) -> uint64_t { return ent[i]; };
      E = group_kth<true>(lget, N, K, kth_all[wv]);
      theta = merge_theta(a, q, E);
      banded = a.band.on;
      kept = compact(lget, E);
    }
  } else {
    E = group_kth<true>(gget, N, K, kth_all[wv]);
    theta = merge_theta(a, q, E);
    if (a.band.on) {
      banded = true;
      int c = 0;
      for (int base = 0; base < N; base += 64) {
        const uint64_t e = base + lane < N ? gget(base + lane) : ~0ull;
        c += __popcll(__builtin_amdgcn_ballot_w64(e != ~0ull && (e <= E || (uint32_t)(e >> 32) < theta)));
      }
      if (c > a.carry_cap) {   // the band overflows the carry: keep its first carry_cap entries
        E = group_kth<true>(gget, N, a.carry_cap, kth_all[wv]);
        theta = theta < (uint32_t)(E >> 32) ? theta : (uint32_t)(E >> 32);
        banded = false;
      }
    }
    kept = compact(gget, E);
  }
  if (lane == 0) {
    if (kept > a.carry_cap) kept = a.carry_cap;
    a.carry_out_cnt[q] = kept;
    a.thr[q] = theta;
    if (kept < K || (!a.band.on && kept != K)) atomicAdd(&a.flags[0], 1);
  }
}

// Threshold-mode merge with the entries in registers (up to 16 per lane, N <= 1024): the LDS
// footprint per wave is the source table and the radix histogram only, so several times more
// waves are resident than with an LDS copy of the entries, and each entry finds its source with
// one LDS read of a per-entry source map instead of a binary search over the source offsets.

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// K-th smallest (1 <= K <= N) of the N distinct entries (key << 32 | row) held by one wave in
// registers: v[u] of lane l is entry l + 64 u, for l + 64 u < N. Radix select over the 32-bit keys
// (8 bits per level, one LDS histogram, the chosen bin broadcast by readlane), so every level costs
// one pass of 32-bit arithmetic and two wave syncs; entries that share the K-th key are resolved
// by the 64-bit select over that key's entries only. hist: >= 4 * 256 uint32 of LDS.
template <int U>
__device__ __forceinline__ uint64_t wave_kth_regs(const uint64_t (&v)[U], int N, int K, uint32_t* hist) {
  const int lane = threadIdx.x & 63;
  uint32_t kmn = ~0u, kmx = 0u;
  uint64_t emn = ~0ull, emx = 0ull;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (lane + 64 * u < N) {
      const uint32_t key = (uint32_t)(v[u] >> 32);
      kmn = min(kmn, key);
      kmx = max(kmx, key);
    }
  kmn = wave_min_u32(kmn);
  kmx = wave_max_u32(kmx);
  if (K >= N || K <= 1) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lane + 64 * u < N) {
        emn = v[u] < emn ? v[u] : emn;
        emx = v[u] > emx ? v[u] : emx;
      }
    return K >= N ? wave_max_u64(emx) : wave_min_u64(emn);
  }
  // four sub-histograms (lane & 3) cut the same-bin conflicts of the LDS atomics by four
#pragma unroll
  for (int g = 0; g < 4; ++g) *(u32x4*)(hist + 256 * g + 4 * lane) = (u32x4){0u, 0u, 0u, 0u};
  wave_sync();
  uint32_t* hsub = hist + 256 * (lane & 3);
  uint32_t base = kmn;
  int below = 0;
  const uint32_t span = kmx - kmn;
  int sh = span ? 32 - __clz((int)span) : 0;
  sh = sh > 8 ? sh - 8 : 0;
  for (;;) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lane + 64 * u < N) {
        const uint32_t key = (uint32_t)(v[u] >> 32);
        const uint32_t r = (key - base) >> sh;
        if (key >= base && r < 256u) atomicAdd(&hsub[r], 1u);
      }
    wave_sync();
    u32x4 c = (u32x4){0u, 0u, 0u, 0u};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      c += *(const u32x4*)(hist + 256 * g + 4 * lane);
      *(u32x4*)(hist + 256 * g + 4 * lane) = (u32x4){0u, 0u, 0u, 0u};
    }
    wave_sync();
    const uint32_t tot = c[0] + c[1] + c[2] + c[3];
    const uint32_t incl = wave_incl_scan_u32(tot, lane);
    const uint32_t need = (uint32_t)(K - below);
    const int fl = __ffsll((long long)__builtin_amdgcn_ballot_w64(incl >= need)) - 1;
    // lane fl holds the bin with the K-th key: its index, the count below it and its count
    uint32_t acc = incl - tot, b = 0, cb = c[0];
    if (acc + c[0] < need) {
      acc += c[0]; b = 1; cb = c[1];
      if (acc + c[1] < need) {
        acc += c[1]; b = 2; cb = c[2];
        if (acc + c[2] < need) { acc += c[2]; b = 3; cb = c[3]; }
      }
    }
    const uint32_t bstar = (uint32_t)__builtin_amdgcn_readlane((int)(4 * lane + b), fl);
    below += __builtin_amdgcn_readlane((int)acc, fl);
    const int cbin = __builtin_amdgcn_readlane((int)cb, fl);
    base += bstar << sh;
    if (below + cbin == K || below + 1 == K || sh == 0) {
      // below + cbin == K: every entry of the bin is kept and the K-th is the bin's largest;
      // below + 1 == K: the K-th is the bin's smallest.
      // sh == 0: the bin is one key; the K-th is the (K - below)-th of its entries by row.
      const int want = K - below;
      uint64_t m2 = 0ull, m1 = ~0ull;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (lane + 64 * u < N) {
          const uint32_t key = (uint32_t)(v[u] >> 32);
          if (key >= base && ((key - base) >> sh) == 0u) {
            m2 = v[u] > m2 ? v[u] : m2;
            m1 = v[u] < m1 ? v[u] : m1;
          }
        }
      if (want == cbin) return wave_max_u64(m2);
      if (want == 1) return wave_min_u64(m1);
      const uint32_t kk = base;
      auto each = [&](auto f) {
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (lane + 64 * u < N && (uint32_t)(v[u] >> 32) == kk) f(v[u]);
      };
      return group_kth_each<true>(each, cbin, want, hist);
    }
    sh = sh > 8 ? sh - 8 : 0;
  }
}

// U = register entries per lane (64 U entries per query): 16, or 32 for the rounds whose union may
// exceed 1024 entries (large shortlists after a x4 growth step).
template <int U>
__global__ __launch_bounds__(256) void merge_reg_kernel(MergeArgs a) {
  __shared__ int off_all[4][MERGE_MAX_SRC + 1];
  __shared__ uint16_t map_all[4][64 * U];
  __shared__ __attribute__((aligned(16))) uint32_t kth_all[4][4 * 256 + 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + wv;
  if (q >= (a.n_dev ? *a.n_dev : a.n)) return;   // wave-uniform; only wave-level sync below
  int* off = off_all[wv];
  uint16_t* map = map_all[wv];
  const int nsrc = 1 + 2 * a.slices;
  const long subq = (long)q * a.slices_max * 2;
  const int P = (nsrc + 63) / 64;   // sources per lane (<= 5)
  int c[5] = {0, 0, 0, 0, 0};
  {
    int tot = 0;
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int sidx = lane * P + u;
      if (u < P && sidx < nsrc) c[u] = sidx == 0 ? a.carry_in_cnt[q] : a.sub_cnt[subq + sidx - 1];
      tot += c[u];
    }
    const int excl = (int)wave_incl_scan_u32((uint32_t)tot, lane) - tot;
    int acc = excl;
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int sidx = lane * P + u;
      if (u < P && sidx < nsrc) off[sidx] = acc;
      acc += c[u];
    }
    if (lane == 63) off[nsrc] = acc;
    wave_sync();
  }
  const int N = off[nsrc];
  const int K = merge_rank(a, a.kc[q]);
  const uint64_t* cin = a.carry_in + (long)q * a.carry_cap;
  const uint64_t* sub = a.sub + subq * a.sub_cap;
  const uint32_t self = a.self_ids ? (uint32_t)a.self_ids[q] : 0xffffffffu;
  auto key_of = [&](int s, int j) -> uint64_t {
    if (s == 0) return cin[j];
    const uint64_t e = sub[(long)(s - 1) * a.sub_cap + j];
    return (uint32_t)e == self ? ~0ull : pool_key(e);
  };
  uint64_t* cout = a.carry_out + (long)q * a.carry_cap;
  uint64_t E = ~0ull;
  uint32_t theta = 0;   // the new threshold (set when N > K)
  bool banded = a.band.on;
  int kept = 0;
  if (N > 64 * U) {
    // large union (only with unusual shortlists or ratios): radix select over global memory
    auto gget = [&](int i) -> uint64_t {
      int lo = 0, hi = nsrc - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
      }
      return key_of(lo, i - off[lo]);
    };
    E = group_kth<true>(gget, N, K, kth_all[wv]);
    theta = merge_theta(a, q, E);
    auto keep_of = [&](uint64_t e) { return e != ~0ull && (e <= E || (a.band.on && (uint32_t)(e >> 32) < theta)); };
    int cnt = 0;
    if (a.band.on)
      for (int base = 0; base < N; base += 64)
        cnt += __popcll(__builtin_amdgcn_ballot_w64(base + lane < N && keep_of(gget(base + lane))));
    if (cnt > a.carry_cap) {   // the band overflows the carry: keep its first carry_cap entries
      E = group_kth<true>(gget, N, a.carry_cap, kth_all[wv]);
      theta = theta < (uint32_t)(E >> 32) ? theta : (uint32_t)(E >> 32);
      banded = false;
    }
    // the entries up to E first, then (banded) the band beyond E
    for (int pass = 0; pass < (banded ? 2 : 1); ++pass) {
      for (int base = 0; base < N; base += 64) {
        const int i = base + lane;
        uint64_t e = 0;
        bool keep = false;
        if (i < N) {
          e = gget(i);
          keep = e != ~0ull && (pass == 0 ? e <= E : (e > E && keep_of(e)));
        }
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        const int pos = kept + __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
        if (keep && pos < a.carry_cap) cout[pos] = e;
        kept += __popcll(km);
      }
    }
  } else {
    // per-entry source map of the sub-pool entries (the carry's entries, slots [0, off[1]), need
    // none: one lane would otherwise write all of them), then all loads of the lane in flight
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int sidx = lane * P + u;
      if (u < P && sidx > 0 && sidx < nsrc) {
        const int o = off[sidx];
        for (int j = 0; j < c[u]; ++j) map[o + j] = (uint16_t)sidx;
      }
    }
    wave_sync();
    const int ncarry = off[1];
    const int used = (N + 63) >> 6;   // slots holding entries (wave-uniform)
    // branch-free: every lane computes a valid address for each slot (idle slots read the first
    // carry entry), so all U loads are in flight together and one wait covers them; the key
    // conversion follows in a second pass
    uint64_t v[U];
    int src[U], jj[U];
    // the map holds 64 U slots and off[0] always exists: both reads are unconditional (slots
    // past N read stale values that the selects discard), so they too are in flight together
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = lane + 64 * u;
      const int m = (int)map[i];
      src[u] = i < N ? (i < ncarry ? 0 : m) : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = off[src[u] < 0 ? 0 : src[u]];
    __builtin_amdgcn_sched_barrier(0);   // every offset read is issued before its first use
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = src[u] < 0 ? 0 : lane + 64 * u - jj[u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = src[u] < 0 ? 0 : src[u];
      const uint64_t* p = s == 0 ? cin + jj[u] : sub + (long)(s - 1) * a.sub_cap + jj[u];
      v[u] = u < used ? *p : ~0ull;   // slots past the union: no load
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t e = v[u];
      v[u] = src[u] < 0 ? ~0ull
             : (src[u] == 0 ? e : ((uint32_t)e == self ? ~0ull : pool_key(e)));
    }
    if (N > K) {
      E = wave_kth_regs<U>(v, N, K, kth_all[wv]);
      theta = merge_theta(a, q, E);
      if (banded) {
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (64 * u >= N) break;   // wave-uniform
          const uint64_t e = v[u];
          cnt += __popcll(__builtin_amdgcn_ballot_w64(lane + 64 * u < N && e != ~0ull &&
                                                      (e <= E || (uint32_t)(e >> 32) < theta)));
        }
        if (cnt > a.carry_cap) {   // the band overflows the carry: keep its first carry_cap entries
          E = wave_kth_regs<U>(v, N, a.carry_cap, kth_all[wv]);
          theta = theta < (uint32_t)(E >> 32) ? theta : (uint32_t)(E >> 32);
          banded = false;
        }
      }
    }
    // the entries up to E first, then (banded) the band beyond E below theta: the rerank takes
    // the carry's first K' entries as the shortlist proper
    auto emit = [&](auto band_pass) {   // band_pass: std::false_type, then std::true_type
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (64 * u >= N) break;   // wave-uniform
        const uint64_t e = v[u];
        const bool keep = lane + 64 * u < N && e != ~0ull &&
                          (band_pass ? (e > E && (uint32_t)(e >> 32) < theta) : e <= E);
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        const int pos = kept + __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
        if (keep && pos < a.carry_cap) cout[pos] = e;
        kept += __popcll(km);
      }
    };
    emit(std::false_type{});
    if (banded && N > K) emit(std::true_type{});
  }
  if (lane == 0) {
    if (kept > a.carry_cap) kept = a.carry_cap;
    a.carry_out_cnt[q] = kept;
    if (N > K) {
      a.thr[q] = theta;
      if (kept < K || (!a.band.on && kept != K)) atomicAdd(&a.flags[0], 1);
    }
  }
}

// ------------------------------------------------------------------------------------------------
// shortlist certificate (one thread per query, after the exact rerank)
// ------------------------------------------------------------------------------------------------
// Every corpus row x outside the shortlist has an approximate key s~(x) >= tau (tau = the key of
// the K'-th entry, thr[q]). The key is s~ = ||r~||^2/2 - q~.r~ accumulated in FP32 from operands
// q~, r~ rounded for the scan, so with alpha = gamma (||q~|| max||r~|| + max||r~||^2 / 2) bounding
// the accumulation error, ||q~ - r~||^2 = 2 s(q~, r~) + ||q~||^2 >= 2 (tau - alpha) + ||q~||^2, and
// ||q - r|| >= ||q~ - r~|| - ||q~ - q|| - max ||r~ - r|| (triangle inequality). If that lower bound
// exceeds D_k, the k-th exact distance of the shortlist, no row outside it can be among the k
// nearest: the query is certified. Otherwise the kernel leaves in thr[q] the smallest key threshold
// that certifies it -- every row with a key at or above it is provably farther than D_k -- for a
// re-search that admits every row below it.

// Every row whose approximate key exceeds the returned value lies provably farther from query q
// than sqrt(dk), dk an exact squared distance of the rerank: the bound above solved for the key.
__device__ double key_beyond(const FinalArgs& a, int q, double dk) {
  const double qn2 = a.qn2[q], rq = a.qrho[q];
  const double rmax = (double)__builtin_bit_cast(float, a.bnd[BND_RHO + a.bslot]);
  const double nmax = (double)__builtin_bit_cast(float, a.bnd[BND_NRM + a.bslot]);
  const double alpha = a.gamma * (__builtin_sqrt(qn2) * nmax + 0.5 * nmax * nmax) * (1.0 + 1e-9);
  const double Dk = __builtin_sqrt(dk) * (1.0 + 1e-12);   // covers the rerank's own rounding
  const double r = Dk + rq + rmax;
  const double tn = 0.5 * (r * r - qn2) + alpha;
  return tn + 1e-12 * (r * r + qn2 + 1.0);                  // FP64 rounding of this expression
}

__device__ void certify_query(const FinalArgs& a, int q, int k, int Kq, double dk) {
  const uint32_t t = a.thr[q];
  uint32_t need = 0;
  if (Kq < k && Kq < a.rows && t != ORD_POS_INF) {
    // rows were filtered out and fewer than k remain (a predicted early threshold set too low):
    // search the query again with every row admitted
    need = ORD_POS_INF;
    atomicAdd(&a.flags[2], 1);
  } else if (Kq >= k && t != ORD_POS_INF) {
    const double tn = key_beyond(a, q, dk);
    const double tau = (double)ord2f(t);
    if (!(tau > tn)) {
      need = f2ord((float)next_up_f32(tn));
      if (need == 0) need = 1;
      atomicAdd(&a.flags[2], 1);
    }
  }
  a.thr[q] = need;
}

// ------------------------------------------------------------------------------------------------
// final kernel: one 256-thread block per query
// ------------------------------------------------------------------------------------------------
constexpr int FINAL_MAX = 2048;

__device__ __forceinline__ void bitonic_sort_lds_pair(uint64_t* k, uint32_t* v, int P) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int t = tid; t < P / 2; t += nt) {
        const int pos = 2 * t - (t & (stride - 1));
        const bool up = ((pos & size) == 0);
        const uint64_t x = k[pos], y = k[pos + stride];
        const uint32_t xv = v[pos], yv = v[pos + stride];
        const bool gt = (x > y) || (x == y && xv > yv);
        if (gt == up) { k[pos] = y; k[pos + stride] = x; v[pos] = yv; v[pos + stride] = xv; }
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void final_block_kernel(FinalArgs a) {
  __shared__ uint64_t ek[FINAL_MAX];
  __shared__ uint32_t ev[FINAL_MAX];
  __shared__ double qrow[4096];
  const int q = blockIdx.x;
  if (q >= (a.n_dev ? *a.n_dev : a.n)) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  int cnt = a.carry_cnt[q];
  if (cnt > FINAL_MAX) cnt = FINAL_MAX;
  int K = a.kc[q];
  const int k = a.kq[q];
  const uint64_t* src = a.carry + (long)q * a.carry_cap;
  // the merge leaves exactly the shortlist in the carry (at most K entries): rerank all of them
  for (int i = tid; i < cnt; i += nt) ek[i] = src[i];
  __syncthreads();
  const int Kq = cnt;   // every carry entry: the shortlist and its certificate band
  // query master row
  const int d = a.d;
  const bool dq = d <= 4096;
  if (dq) {
    for (int j = tid; j < d; j += nt)
      qrow[j] = a.master_f64 ? ((const double*)a.Qm)[(long)q * a.ldqm + j]
                             : (double)((const float*)a.Qm)[(long)q * a.ldqm + j];
  }
  __syncthreads();
  // exact rerank: canonical distance, sequential FMA over the dims
  for (int t = tid; t < Kq; t += nt) {
    const uint32_t idx = (uint32_t)(ek[t] & 0xffffffffu);
    double acc = 0.0;
    if (a.master_f64) {
      const double* r = (const double*)a.Xm + (long)idx * a.ldxm;
      const double* qq = (const double*)a.Qm + (long)q * a.ldqm;
      for (int j = 0; j < d; ++j) {
        const double t1 = (dq ? qrow[j] : qq[j]) - r[j];
        acc = __builtin_fma(t1, t1, acc);
      }
    } else {
      const float* r = (const float*)a.Xm + (long)idx * a.ldxm;
      const float* qq = (const float*)a.Qm + (long)q * a.ldqm;
      int j = 0;
      if ((a.ldxm & 3) == 0) {
        for (; j + 4 <= d; j += 4) {
          const f32x4 rv = *(const f32x4*)(r + j);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double t1 = (dq ? qrow[j + u] : (double)qq[j + u]) - (double)rv[u];
            acc = __builtin_fma(t1, t1, acc);
          }
        }
      }
      for (; j < d; ++j) {
        const double t1 = (dq ? qrow[j] : (double)qq[j]) - (double)r[j];
        acc = __builtin_fma(t1, t1, acc);
      }
    }
    ev[t] = idx;
    ek[t] = __builtin_bit_cast(uint64_t, acc);   // acc >= +0: bit order == numeric order
  }
  int P2 = 1;
  while (P2 < Kq) P2 <<= 1;
  __syncthreads();
  for (int i = Kq + tid; i < P2; i += nt) { ek[i] = ~0ull; ev[i] = ~0u; }
  bitonic_sort_lds_pair(ek, ev, P2);
  if (a.certify && tid == 0) certify_query(a, q, k, Kq, Kq >= k ? __builtin_bit_cast(double, ek[k - 1]) : 0.0);
  const long o = a.out_off[q];
  for (int t = tid; t < k; t += nt) {
    double dist;
    int32_t idx;
    if (t < Kq) { dist = __builtin_bit_cast(double, ek[t]); idx = (int32_t)ev[t]; }
    else { dist = __builtin_inf(); idx = -1; }
    if (a.out_f64) ((double*)a.out_dist)[o + t] = dist;
    else ((float*)a.out_dist)[o + t] = (float)dist;
    a.out_idx[o + t] = idx;
  }
}

// Final rerank, one wave per query (shortlists up to FW_MAX). Each lane reranks up to four
// candidates at a time: four independent FMA chains, each in the canonical sequential order.
constexpr int FW_MAX = 1024;

__device__ __forceinline__ void ce_pair(uint64_t& xk, uint32_t& xv, uint64_t& yk, uint32_t& yv, bool up) {
  const bool gt = (xk > yk) || (xk == yk && xv > yv);
  const bool sw = gt == up;
  const uint64_t lk = sw ? yk : xk;
  const uint32_t lv = sw ? yv : xv;
  yk = sw ? xk : yk;
  yv = sw ? xv : yv;
  xk = lk;
  xv = lv;
}

// Ascending wave bitonic sort of P (a power of two, 64 <= P <= 512) (key, value) pairs in LDS by
// (key, value). Stages of stride >= 8 run through LDS; every stage of stride < 8 runs in registers
// on 8-entry blocks (keys and values read and written as 16-byte units).
__device__ void wave_bitonic_pair_r8(uint64_t* k, uint32_t* v, int P) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63;
  const int nblk = P >> 3;
  auto load8 = [&](int b, uint64_t (&rk)[8], uint32_t (&rv)[8]) {
    const u64x2* pk = (const u64x2*)(k + 8 * b);
    const u32x4* pv = (const u32x4*)(v + 8 * b);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u64x2 t = pk[i];
      rk[2 * i] = t[0];
      rk[2 * i + 1] = t[1];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const u32x4 t = pv[i];
      rv[4 * i] = t[0]; rv[4 * i + 1] = t[1]; rv[4 * i + 2] = t[2]; rv[4 * i + 3] = t[3];
    }
  };
  auto store8 = [&](int b, const uint64_t (&rk)[8], const uint32_t (&rv)[8]) {
    u64x2* pk = (u64x2*)(k + 8 * b);
    u32x4* pv = (u32x4*)(v + 8 * b);
#pragma unroll
    for (int i = 0; i < 4; ++i) pk[i] = (u64x2){rk[2 * i], rk[2 * i + 1]};
#pragma unroll
    for (int i = 0; i < 2; ++i) pv[i] = (u32x4){rv[4 * i], rv[4 * i + 1], rv[4 * i + 2], rv[4 * i + 3]};
  };
  auto ce = [&](uint64_t (&rk)[8], uint32_t (&rv)[8], int i, int j, bool up) {
    ce_pair(rk[i], rv[i], rk[j], rv[j], up);
  };
  auto tail3 = [&](uint64_t (&rk)[8], uint32_t (&rv)[8], bool up) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ce(rk, rv, i, i + 4, up);
#pragma unroll
    for (int i : {0, 1, 4, 5}) ce(rk, rv, i, i + 2, up);
#pragma unroll
    for (int i : {0, 2, 4, 6}) ce(rk, rv, i, i + 1, up);
  };
  wave_sync();
  for (int b = lane; b < nblk; b += 64) {
    uint64_t rk[8];
    uint32_t rv[8];
    load8(b, rk, rv);
    ce(rk, rv, 0, 1, true); ce(rk, rv, 2, 3, false); ce(rk, rv, 4, 5, true); ce(rk, rv, 6, 7, false);
    ce(rk, rv, 0, 2, true); ce(rk, rv, 1, 3, true); ce(rk, rv, 4, 6, false); ce(rk, rv, 5, 7, false);
    ce(rk, rv, 0, 1, true); ce(rk, rv, 2, 3, true); ce(rk, rv, 4, 5, false); ce(rk, rv, 6, 7, false);
    tail3(rk, rv, (b & 1) == 0);
    store8(b, rk, rv);
  }
  for (int size = 16; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride >= 8; stride >>= 1) {
      wave_sync();
      for (int t = lane; t < P / 2; t += 64) {
        const int pos = 2 * t - (t & (stride - 1));
        const bool up = ((pos & size) == 0);
        const uint64_t x = k[pos], y = k[pos + stride];
        const uint32_t xv = v[pos], yv = v[pos + stride];
        const bool gt = (x > y) || (x == y && xv > yv);
        if (gt == up) { k[pos] = y; k[pos + stride] = x; v[pos] = yv; v[pos + stride] = xv; }
      }
    }
    wave_sync();
    for (int b = lane; b < nblk; b += 64) {
      uint64_t rk[8];
      uint32_t rv[8];
      load8(b, rk, rv);
      tail3(rk, rv, ((8 * b) & size) == 0);
      store8(b, rk, rv);
    }
  }
  wave_sync();
}

constexpr int FW_COOP_D = 256;   // widest rows of the cooperative gather (query row kept in LDS)

// COOP: FP32 master rows of 16-byte aligned stride, d a multiple of 16 up to FW_COOP_D (the
// cooperative gather below); otherwise the per-lane gather. Two instantiations keep each gather's
// registers apart.
__host__ __device__ inline bool final_coop_ok(const FinalArgs& a) {
  return !a.master_f64 && (a.ldxm & 3) == 0 && (a.ldqm & 3) == 0 && (a.d & 15) == 0 && a.d <= FW_COOP_D;
}

template <bool COOP>
__global__ __launch_bounds__(256) void final_wave_kernel(FinalArgs a) {
  __shared__ __attribute__((aligned(16))) uint64_t dk_all[4][FW_MAX];
  __shared__ __attribute__((aligned(16))) uint32_t di_all[4][FW_MAX];
  __shared__ f32x4 tile_all[4][COOP ? 64 * 4 : 1];            // 64 rows x 16 dims per wave
  __shared__ f32x4 qs_all[4][COOP ? FW_COOP_D / 4 : 1];       // the query row
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + wv;
  if (q >= (a.n_dev ? *a.n_dev : a.n)) return;   // wave-uniform
  uint64_t* dk = dk_all[wv];
  uint32_t* di = di_all[wv];
  f32x4* tile = tile_all[wv];
  f32x4* qs = qs_all[wv];
  int cnt = a.carry_cnt[q];
  if (cnt > FW_MAX) cnt = FW_MAX;
  const int K = a.kc[q];
  const int k = a.kq[q];
  const int d = a.d;
  const uint64_t* src = a.carry + (long)q * a.carry_cap;
  constexpr bool coop = COOP;
  if (coop) {
    const f32x4* qr = (const f32x4*)((const float*)a.Qm + (long)q * a.ldqm);
    for (int i = lane; i < d / 4; i += 64) qs[i] = qr[i];
  }
  auto sort_first = [&](int count) {   // sorts dk / di [0, count) by (distance, row)
    int P = 64;
    while (P < count) P <<= 1;
    wave_sync();
    for (int i = count + lane; i < P; i += 64) { dk[i] = ~0ull; di[i] = ~0u; }
    wave_bitonic_pair_r8(dk, di, P);   // 64 <= P <= FW_MAX
  };
  // The merges write the carry's K' smallest keys first and the certificate band after them.
  // Phase 1 reranks the first min(K', cnt) entries; the k-th of their exact distances bounds the
  // carry's k-th from above (it is the k-th of a subset), so a band entry whose key proves it
  // farther than that bound cannot be among the k nearest, and phase 2 reranks only the others.
  // (The pruning needs no order in the carry; the order only makes the bound tight.) One copy of
  // the gather serves both phases: list[0, count) -> dk / di [dst, dst + count), list being the
  // carry (phase 1) or this wave's LDS (phase 2: each slot is read, then written, by one lane).
  const int Kc = K < cnt ? K : cnt;
  const uint64_t* list = src;
  int count = Kc, dst = 0, nr = Kc;
  bool sorted = false;   // dk / di [0, nr) already in (distance, row) order
  // the rounding radii of the first 256 band entries' rows, loaded before phase 1's gather
  const bool row_radii = a.prune && a.rrho && cnt > Kc;
  float rxp[4] = {0.f, 0.f, 0.f, 0.f};
  if (row_radii) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = Kc + 64 * u + lane;
      if (t < cnt) rxp[u] = a.rrho[(uint32_t)(src[t] & 0xffffffffu)];
    }
  }
  for (int phase = 0;; ++phase) {
    // Cooperative gather: lanes 4j .. 4j+3 load the four 16-byte parts of one row's 64-byte chunk
    // (16 dims), so a load instruction covers 16 rows and every sector is requested by one
    // instruction only (per-lane row loads touch each 64-byte sector from four instructions, three
    // of which wait for the first one's miss). Each chunk goes through this wave's LDS tile (parts
    // swizzled by row so that the per-row reads are free of bank conflicts) to the lane that owns
    // the row, whose FMA chain takes the dims in order: the same canonical sum as the per-lane form.
    for (int b = 0; coop && b < count; b += 64) {
      constexpr int FC = 8;                            // chunks in flight: 32 loads per lane
      const int slot = lane >> 2, part = lane & 3;
      const float* rp[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = b + 16 * t + slot;
        const uint32_t row = i < count ? (uint32_t)(list[i] & 0xffffffffu) : 0u;   // row 0 for idle slots
        rp[t] = (const float*)a.Xm + (long)row * a.ldxm + 4 * part;
      }
      const int mine = b + lane;
      const bool okm = mine < count;
      const uint32_t myrow = okm ? (uint32_t)(list[mine] & 0xffffffffu) : 0u;
      double acc = 0.0;
      for (int c0 = 0; c0 < d; c0 += 16 * FC) {
        f32x4 v[FC][4];
#pragma unroll
        for (int c = 0; c < FC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (c0 + 16 * c < d) v[c][t] = *(const f32x4*)(rp[t] + c0 + 16 * c);
#pragma unroll
        for (int c = 0; c < FC; ++c) {
          if (c0 + 16 * c >= d) break;   // wave-uniform
          wave_sync();                   // the previous chunk's reads of the tile are done
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int r = 16 * t + slot;
            tile[r * 4 + (part ^ ((r >> 2) & 3))] = v[c][t];
          }
          wave_sync();
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            const f32x4 m = tile[lane * 4 + (p ^ ((lane >> 2) & 3))];
            const f32x4 qv = qs[(c0 + 16 * c) / 4 + p];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const double t1 = (double)qv[e] - (double)m[e];
              acc = __builtin_fma(t1, t1, acc);
            }
          }
        }
      }
      if (okm) { dk[dst + mine] = __builtin_bit_cast(uint64_t, acc); di[dst + mine] = myrow; }
    }
      for (int b = 0; !coop && b < count; b += 256) {
        uint32_t idx[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = b + u * 64 + lane;
          ok[u] = t < count;
          idx[u] = ok[u] ? (uint32_t)(list[t] & 0xffffffffu) : 0u;   // row 0 stands in for idle lanes
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (!a.master_f64) {
          const float* qr = (const float*)a.Qm + (long)q * a.ldqm;
          const float* r[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) r[u] = (const float*)a.Xm + (long)idx[u] * a.ldxm;
          int j = 0;
          if ((a.ldxm & 3) == 0 && (a.ldqm & 3) == 0) {
            // 8 x 16 bytes per row per step: 32 independent loads in flight per lane (the gather
            // is latency-bound, rows are scattered over the corpus)
            constexpr int U = 8;
            for (; j + 4 * U <= d; j += 4 * U) {
              f32x4 qv[U], rv[U][4];
#pragma unroll
              for (int w = 0; w < U; ++w) {
                qv[w] = *(const f32x4*)(qr + j + 4 * w);
#pragma unroll
                for (int u = 0; u < 4; ++u) rv[w][u] = *(const f32x4*)(r[u] + j + 4 * w);
              }
#pragma unroll
              for (int w = 0; w < U; ++w)
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const double t1 = (double)qv[w][e] - (double)rv[w][u][e];
                    acc[u] = __builtin_fma(t1, t1, acc[u]);
                  }
            }
            for (; j + 4 <= d; j += 4) {
              const f32x4 qv = *(const f32x4*)(qr + j);
              f32x4 rv[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) rv[u] = *(const f32x4*)(r[u] + j);
#pragma unroll
              for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const double t1 = (double)qv[e] - (double)rv[u][e];
                  acc[u] = __builtin_fma(t1, t1, acc[u]);
                }
            }
          }
          for (; j < d; ++j) {
            const double qj = (double)qr[j];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const double t1 = qj - (double)r[u][j];
              acc[u] = __builtin_fma(t1, t1, acc[u]);
            }
          }
        } else {
          const double* qr = (const double*)a.Qm + (long)q * a.ldqm;
          const double* r[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) r[u] = (const double*)a.Xm + (long)idx[u] * a.ldxm;
          for (int j = 0; j < d; ++j) {
            const double qj = qr[j];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const double t1 = qj - r[u][j];
              acc[u] = __builtin_fma(t1, t1, acc[u]);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = b + u * 64 + lane;
          if (ok[u]) { dk[dst + t] = __builtin_bit_cast(uint64_t, acc[u]); di[dst + t] = idx[u]; }
        }
      }
    if (phase == 1 || cnt <= Kc) break;
    const bool prune = a.prune && Kc >= k;
    double dk1 = 0.0, rq = 0.0, qn2 = 0.0, alpha = 0.0, rmax = 0.0;
    if (prune) {
      sort_first(Kc);
      sorted = true;
      // the certificate's bound (key_beyond) with the entry's own row radius when known
      qn2 = a.qn2[q];
      rq = a.qrho[q];
      rmax = (double)__builtin_bit_cast(float, a.bnd[BND_RHO + a.bslot]);
      const double nmax = (double)__builtin_bit_cast(float, a.bnd[BND_NRM + a.bslot]);
      alpha = a.gamma * (__builtin_sqrt(qn2) * nmax + 0.5 * nmax * nmax) * (1.0 + 1e-9);
      dk1 = __builtin_sqrt(__builtin_bit_cast(double, dk[k - 1])) * (1.0 + 1e-12);
    }
    // a band entry stays unless its key proves its row farther than the k-th of phase 1
    auto keep_entry = [&](uint64_t e, double rx) -> bool {
      if (!prune) return true;
      const double r = dk1 + rq + rx;
      const double tn = 0.5 * (r * r - qn2) + alpha + 1e-12 * (r * r + qn2 + 1.0);
      return (double)ord2f((uint32_t)(e >> 32)) <= tn;
    };
    auto emit = [&](int b, bool keep, uint64_t e) {
      const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
      const int pos = nr + __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
      if (keep) dk[pos] = e;
      nr += __popcll(km);
    };
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = Kc + 64 * u;
      if (b >= cnt) break;   // wave-uniform
      const int t = b + lane;
      const uint64_t e = t < cnt ? src[t] : ~0ull;
      emit(b, t < cnt && keep_entry(e, row_radii ? (double)rxp[u] : rmax), e);
    }
    for (int b = Kc + 256; b < cnt; b += 64) {
      const int t = b + lane;
      const uint64_t e = t < cnt ? src[t] : ~0ull;
      const double rx = row_radii && t < cnt ? (double)a.rrho[(uint32_t)(e & 0xffffffffu)] : rmax;
      emit(b, t < cnt && keep_entry(e, rx), e);
    }
    wave_sync();
    if (nr == Kc) break;   // every band entry pruned: the first Kc are the answer, sorted
    sorted = false;
    list = dk + Kc;
    count = nr - Kc;
    dst = Kc;
  }
  if (!sorted) sort_first(nr);
  if (a.certify && lane == 0) certify_query(a, q, k, nr, nr >= k ? __builtin_bit_cast(double, dk[k - 1]) : 0.0);
  const long o = a.out_off[q];
  for (int t = lane; t < k; t += 64) {
    double dist;
    int32_t id;
    if (t < nr) { dist = __builtin_bit_cast(double, dk[t]); id = (int32_t)di[t]; }
    else { dist = __builtin_inf(); id = -1; }
    if (a.out_f64) ((double*)a.out_dist)[o + t] = dist;
    else ((float*)a.out_dist)[o + t] = (float)dist;
    a.out_idx[o + t] = id;
  }
}

// ------------------------------------------------------------------------------------------------
// per-query control arrays for a uniform-k search (no host upload): thresholds (+inf for real
// queries, -inf for padding queries so nothing survives), shortlist, k, output offset, self id
// ------------------------------------------------------------------------------------------------
// Per-query control state. n_dev (optional) holds the real query count on the device, overriding
// n: a search captured in a graph for n_pad query slots then serves any count up to n_pad.
// write_k = 0 leaves kc / kq / off to the caller (per-query k copied from the host).
// write_thr = 0 keeps the real queries' thresholds (preset by the caller: a re-search).
__global__ void init_queries_kernel(uint32_t* thr, int* kc, int* kq, long* off, int* self_ids,
                                    int* carry_cnt, long n, long n_pad, int k, int shortlist,
                                    long self_offset, long m, const int* n_dev, int write_k, int write_thr) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  const long nr = n_dev ? (long)*n_dev : n;
  if (i >= nr) thr[i] = ORD_NEG_INF;
  else if (write_thr) thr[i] = ORD_POS_INF;
  if (write_k) kc[i] = shortlist;
  if (self_ids) {
    const long r = self_offset + i;
    self_ids[i] = (i < nr && r >= 0 && r < m) ? (int)r : -1;
  }
  if (i < nr) {
    if (write_k) {
      kq[i] = k;
      off[i] = i * (long)k;
    }
    carry_cnt[i] = 0;
  }
}

hipError_t launch_init_queries(uint32_t* thr, int* kc, int* kq, long* off, int* self_ids,
                               int* carry_cnt, long n, long n_pad, int k, int shortlist,
                               long self_offset, long m, const int* n_dev, int write_k,
                               int write_thr, hipStream_t stream) {
  if (n_pad <= 0) return hipSuccess;
  hipLaunchKernelGGL(init_queries_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0, stream,
                     thr, kc, kq, off, self_ids, carry_cnt, n, n_pad, k, shortlist, self_offset, m,
                     n_dev, write_k, write_thr);
  return hipGetLastError();
}

// one thread per 4-byte word of the gathered rows
__global__ void gather_rows_kernel(uint32_t* dst, const uint32_t* src, const int* idx, long n, long words) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * words) return;
  const long i = t / words, w = t % words;
  dst[t] = src[(long)idx[i] * words + w];
}

hipError_t launch_gather_rows(void* dst, const void* src, const int* idx, long n, long row_bytes,
                              hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (row_bytes % 4) return hipErrorInvalidValue;
  const long words = row_bytes / 4, total = n * words;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                     (uint32_t*)dst, (const uint32_t*)src, idx, n, words);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// fills
// ------------------------------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_i32_kernel(int* p, int v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
hipError_t launch_pack_fp8(const PackArgs& a, hipStream_t stream) {
  const long blocks = (a.rows_pad + 3) / 4;
  if (blocks <= 0) return hipSuccess;
  if (!a.x8 || a.dp8 % 64 || a.ld8 != a.dp8 + 16 || a.d > a.dp8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pack_fp8_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_pack(const PackArgs& a, hipStream_t stream) {
  const long blocks = (a.rows_pad + 3) / 4;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// every scan: one dimension, wg_rows x slices items rounded up to a multiple of 8 workgroups (the
// item order of ScanArgs::wg_rows)
static hipError_t scan_grid(void (*kern)(ScanArgs), const ScanArgs& a, int groups, hipStream_t stream) {
  if (groups % 4 || a.wg_rows != groups / 4 || a.slices < 1) return hipErrorInvalidValue;
  const long items = (long)a.wg_rows * a.slices;
  if (items > 0x7fffff00L) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kern, dim3((unsigned)((items + 7) / 8 * 8)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <int PREC, int QB, int DC, int SR>
static hipError_t scan_one(const ScanArgs& a, int groups, hipStream_t stream) {
  return scan_grid(scan_kernel<PREC, QB, DC, SR>, a, groups, stream);
}

bool scan_supported(int prec, int qb, int dc) {
  if (prec != PREC_F32 && prec != PREC_BF16) return false;
  if (qb != 1 && qb != 2) return false;
  return dc == 32 || dc == 64 || dc == 96 || dc == 128;
}

// Query blocks of 32 per wave that the scan is built for: 1 and 2 everywhere; 3 and 4 for the FP8
// scan of 64- and 128-byte rows; 3 and 4 for the single-chunk BF16 scan of 64- and 128-dim rows.
// The pipelined threshold scan fits 3 blocks in two resident workgroups per CU for rows of up to
// 8 units (FP8 rows of 64 and 128 bytes, BF16 rows of 32 and 64 dims); wider rows with 3 or 4
// blocks (scan_swz_kernel, 400+ registers per lane) leave one workgroup per CU.
static int scan_units(int prec, int dc) { return prec == PREC_FP8 ? dc / 16 : dc / 8; }
static bool scan_pipe_ok(int prec, int dp, int dc, int qb) {
  if (!(prec == PREC_FP8 || (prec == PREC_BF16 && dp == dc && dc != 96))) return false;
  return qb == 1 || qb == 2 || (qb == 3 && (scan_units(prec, dc) <= 8 || prec == PREC_BF16)) ||
         (qb == 4 && scan_units(prec, dc) <= 8 && (prec == PREC_FP8 || dc == 64));
}
bool scan_qb_ok(int prec, int dp, int dc, int qb) {
  if (qb == 1 || qb == 2) return true;
  if ((qb == 3 || qb == 4) && prec == PREC_FP8) return dc == 64 || dc == 128;
  return (qb == 3 || qb == 4) && prec == PREC_BF16 && dp == dc && (dc == 64 || dc == 128);
}
int scan_wgs_per_cu(int prec, int dp, int dc, int qb) {
  return qb <= 2 || scan_pipe_ok(prec, dp, dc, qb) ? 2 : 1;
}

// Row granularity of the stages (64 or 128 rows): slice sizes are multiples of it. A slice may
// start at any row; the corpus carries 128 rows past round_up(m, 128), so the last stage of a slice
// (whose rows past row_end are masked) never reads past the allocation. Every row stride is a
// multiple of 16 bytes, so the 16-byte DMA loads stay aligned from any starting row.
int scan_stage_rows(int prec, int dp, int dc) {
  return ((prec == PREC_BF16 && dp == dc) || prec == PREC_FP8) ? 128 : 64;
}

hipError_t launch_scan(int prec, int qb, int dc, const ScanArgs& a, int groups, hipStream_t stream) {
  const int sr = scan_stage_rows(prec, a.dp, dc);
  if (a.row_begin < 0 || (a.slice_rows % sr)) return hipErrorInvalidValue;
  if (a.row_end > 0x7fffff80L) return hipErrorInvalidValue;   // int32 row indices
  if (!a.dense && scan_pipe_ok(prec, a.dp, dc, qb)) {
    if (prec == PREC_FP8 && a.dp != dc) return hipErrorInvalidValue;
    const int su = scan_units(prec, dc);
    const bool f8 = prec == PREC_FP8;
#define KNNS_PIPE_CASE(B, S, F) \
    if (qb == B && su == S && f8 == F) return scan_grid(scan_pipe_kernel<B, S, F>, a, groups, stream);
    KNNS_PIPE_CASE(1, 4, false) KNNS_PIPE_CASE(1, 8, false) KNNS_PIPE_CASE(1, 16, false)
    KNNS_PIPE_CASE(2, 4, false) KNNS_PIPE_CASE(2, 8, false) KNNS_PIPE_CASE(2, 16, false)
    KNNS_PIPE_CASE(1, 4, true) KNNS_PIPE_CASE(1, 8, true) KNNS_PIPE_CASE(1, 16, true)
    KNNS_PIPE_CASE(2, 4, true) KNNS_PIPE_CASE(2, 8, true) KNNS_PIPE_CASE(2, 16, true)
    KNNS_PIPE_CASE(3, 4, true) KNNS_PIPE_CASE(3, 8, true) KNNS_PIPE_CASE(3, 4, false) KNNS_PIPE_CASE(3, 8, false)
    KNNS_PIPE_CASE(3, 16, false) KNNS_PIPE_CASE(4, 4, true) KNNS_PIPE_CASE(4, 8, true)
    KNNS_PIPE_CASE(4, 8, false)
#undef KNNS_PIPE_CASE
    return hipErrorInvalidValue;
  }
  if (prec == PREC_FP8) {
    // dc = dp = FP8 row bytes (64, 128, 256); ScanArgs strides are in bytes
    if (a.dp != dc) return hipErrorInvalidValue;
#define KNNS_F8_CASE(B, D) \
    if (qb == B && dc == D) return scan_grid(scan_swz_kernel<B, D / 16, true>, a, groups, stream);
    KNNS_F8_CASE(1, 64) KNNS_F8_CASE(1, 128) KNNS_F8_CASE(1, 256)
    KNNS_F8_CASE(2, 64) KNNS_F8_CASE(2, 128) KNNS_F8_CASE(2, 256)
    KNNS_F8_CASE(3, 64) KNNS_F8_CASE(3, 128) KNNS_F8_CASE(4, 64) KNNS_F8_CASE(4, 128)
#undef KNNS_F8_CASE
    return hipErrorInvalidValue;
  }
  if (prec == PREC_BF16 && a.dp == dc && dc != 96) {
#define KNNS_SWZ_CASE(B, D) \
    if (qb == B && dc == D) return scan_grid(scan_swz_kernel<B, D / 8, false>, a, groups, stream);
    KNNS_SWZ_CASE(1, 32) KNNS_SWZ_CASE(1, 64) KNNS_SWZ_CASE(1, 128)
    KNNS_SWZ_CASE(2, 32) KNNS_SWZ_CASE(2, 64) KNNS_SWZ_CASE(2, 128)
    KNNS_SWZ_CASE(3, 64) KNNS_SWZ_CASE(3, 128) KNNS_SWZ_CASE(4, 64) KNNS_SWZ_CASE(4, 128)
#undef KNNS_SWZ_CASE
    return hipErrorInvalidValue;
  }
#define KNNS_SCAN_CASE(P, B, D, S) \
  if (prec == P && qb == B && dc == D && sr == S) return scan_one<P, B, D, S>(a, groups, stream);
#define KNNS_SCAN_DC(P, B, S) \
  KNNS_SCAN_CASE(P, B, 32, S) KNNS_SCAN_CASE(P, B, 64, S) KNNS_SCAN_CASE(P, B, 96, S) KNNS_SCAN_CASE(P, B, 128, S)
  KNNS_SCAN_DC(PREC_F32, 1, 64) KNNS_SCAN_DC(PREC_F32, 2, 64)
  KNNS_SCAN_DC(PREC_BF16, 1, 64) KNNS_SCAN_DC(PREC_BF16, 2, 64)
  KNNS_SCAN_DC(PREC_BF16, 1, 128) KNNS_SCAN_DC(PREC_BF16, 2, 128)
#undef KNNS_SCAN_DC
#undef KNNS_SCAN_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_merge(const MergeArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  if (1 + 2 * a.slices > MERGE_MAX_SRC) return hipErrorInvalidValue;
  if (a.bitonic)
    hipLaunchKernelGGL(merge_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  else if (a.wide)
    hipLaunchKernelGGL(merge_reg_kernel<32>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL(merge_reg_kernel<16>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_final(const FinalArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  if (a.carry_cap > FINAL_MAX) return hipErrorInvalidValue;
  if (a.carry_cap > FW_MAX) {
    hipLaunchKernelGGL(final_block_kernel, dim3(a.n), dim3(256), 0, stream, a);
  } else if (final_coop_ok(a)) {
    hipLaunchKernelGGL(final_wave_kernel<true>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  } else {
    hipLaunchKernelGGL(final_wave_kernel<false>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

}  // namespace knns

// phem_split.h -- kernels shared by the MST and CC solves: a one-block
// bitonic sort of a sample, flattening of the parent forest, and the bitmap of
// the giant component (the component holding most sampled vertices).
#pragma once
#include "phem_dev.h"

namespace phem {
namespace {

constexpr int SPLIT_BLK = 256;

constexpr int NSAMP = 4096;  // threshold sample size
constexpr int SBLK = 1024;   // one block sorts the sample

constexpr int GSAMP = 1024;  // giant-root sample size

// In-LDS bitonic sort of N keys (a power of two) by NT threads.
template <int N = NSAMP, int NT = SBLK>
__device__ __forceinline__ void bitonic_sort_lds(ull *s) {
  for (int k = 2; k <= N; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < N; i += NT) {
        int ixj = i ^ j;
        if (ixj > i) {
          ull x = s[i], y = s[ixj];
          bool up = (i & k) == 0;
          if ((x > y) == up) {
            s[i] = y;
            s[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// par[x] = root(x) for every vertex.
__global__ __launch_bounds__(SPLIT_BLK) void k_flatten(uint32_t n, uint32_t *par) {
  for (uint64_t i = (uint64_t)blockIdx.x * SPLIT_BLK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * SPLIT_BLK) {
    uint32_t r = uf_root(par, (uint32_t)i);
    if (r != i) par[i] = r;
  }
}

// The most frequent value of the sorted s[0..N) (the smallest on ties),
// ignoring the padding value ~0 (~0 when s holds padding only); every run
// start finds its run's end by binary search.  All NT threads call; thread 0
// receives the result, and its count in *len when len is not NULL.
template <int N, int NT>
__device__ __forceinline__ ull longest_run(const ull *s, ull *len = nullptr) {
  ull best_len = 0, best_val = ~0ull;
  for (int j = threadIdx.x; j < N; j += NT) {
    if (s[j] == ~0ull || (j > 0 && s[j] == s[j - 1])) continue;
    int lo = j, hi = N;  // first index > j with a different value
    while (lo + 1 < hi) {
      const int mid = (lo + hi) >> 1;
      if (s[mid] == s[j]) lo = mid; else hi = mid;
    }
    const ull len = (ull)(hi - j);
    if (len > best_len || (len == best_len && s[j] < best_val)) {
      best_len = len;
      best_val = s[j];
    }
  }
  for (int d = 1; d < WAVE; d <<= 1) {
    const ull ol = __shfl_xor(best_len, d), ov = __shfl_xor(best_val, d);
    if (ol > best_len || (ol == best_len && ov < best_val)) {
      best_len = ol;
      best_val = ov;
    }
  }
  __shared__ ull s_len[NT / WAVE], s_val[NT / WAVE];
  if (lane_id() == 0) {
    s_len[threadIdx.x / WAVE] = best_len;
    s_val[threadIdx.x / WAVE] = best_val;
  }
  __syncthreads();
  ull bl = 0, bv = ~0ull;
  if (threadIdx.x == 0) {
    for (int w = 0; w < NT / WAVE; w++)
      if (s_len[w] > bl || (s_len[w] == bl && s_val[w] < bv)) {
        bl = s_len[w];
        bv = s_val[w];
      }
    if (len) *len = bl;
  }
  __syncthreads();
  return bv;
}

// The most frequent root among GSAMP strided vertices (after k_flatten);
// with count_out, also how many of the GSAMP vertices have it.
__global__ __launch_bounds__(SBLK) void k_pick_giant(uint32_t n,
                                                     const uint32_t *par,
                                                     uint32_t *giant_out,
                                                     uint32_t *count_out) {
  __shared__ ull s[GSAMP];
  for (int j = threadIdx.x; j < GSAMP; j += SBLK)
    s[j] = par[(uint64_t)j * n / GSAMP];
  __syncthreads();
  bitonic_sort_lds<GSAMP>(s);
  ull len = 0;
  const ull r = longest_run<GSAMP, SBLK>(s, &len);
  if (threadIdx.x == 0) {
    *giant_out = (uint32_t)r;
    if (count_out) *count_out = (uint32_t)len;
  }
}

// bit x of `bits` = (par[x] == giant root); one ballot per 64 vertices.
__global__ __launch_bounds__(SPLIT_BLK) void k_giant_bits(uint32_t n,
                                                    const uint32_t *par,
                                                    const uint32_t *giant,
                                                    uint32_t *bits) {
  const uint32_t g = *giant;
  const uint64_t words = ((uint64_t)n + 31) / 32;
  for (uint64_t base = (uint64_t)blockIdx.x * SPLIT_BLK; base < n;
       base += (uint64_t)gridDim.x * SPLIT_BLK) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t mask = __ballot(i < n && par[i] == g);
    if (lane_id() == 0) {
      const uint64_t w = i >> 5;  // i is a multiple of 64 for lane 0
      if (w < words) bits[w] = (uint32_t)mask;
      if (w + 1 < words) bits[w + 1] = (uint32_t)(mask >> 32);
    }
  }
}

// Summary of the giant bitmap kept in LDS by the kernels that test it: bit
// w of the summary = bitmap word w is all ones (its 32 vertices are all in
// the giant component).  Summary word k covers vertices [1024 k, 1024 k +
// 1024); the array holds up to GSUM_MAX words (every vertex id below 2^24)
// and a kernel keeps the first W of them in LDS (W * 4 bytes).  A set bit answers the test without an L2
// request; a clear bit, or a vertex beyond the kernel's W words, falls back
// to the bitmap.
constexpr uint32_t GSUM_MAX = 16384;

__host__ __device__ constexpr uint32_t gsum_words(uint32_t n, uint32_t W) {
  return ((uint64_t)n + 1023) / 1024 < W ? (uint32_t)(((uint64_t)n + 1023) / 1024)
                                         : W;
}

__global__ __launch_bounds__(SPLIT_BLK) void k_giant_summary(
    uint32_t n, const uint32_t *__restrict__ bits, uint32_t *__restrict__ sum) {
  const uint64_t words = ((uint64_t)n + 31) / 32;
  const uint64_t cover = 32ull * gsum_words(n, GSUM_MAX);  // bitmap words covered
  for (uint64_t base = (uint64_t)blockIdx.x * SPLIT_BLK; base < cover;
       base += (uint64_t)gridDim.x * SPLIT_BLK) {
    const uint64_t w = base + threadIdx.x;
    // a partial last word has clear tail bits, so it never reads as full
    const uint64_t mask = __ballot(w < words && bits[w] == 0xFFFFFFFFu);
    if (lane_id() == 0) {
      const uint64_t k = w >> 5;  // w is a multiple of 64 for lane 0
      if (k < cover / 32) sum[k] = (uint32_t)mask;
      if (k + 1 < cover / 32) sum[k + 1] = (uint32_t)(mask >> 32);
    }
  }
}

// Loads the first W summary words into LDS (every thread of the block
// calls it); returns the number of vertices they cover.
template <int NT, uint32_t W>
__device__ __forceinline__ uint32_t gsum_load(uint32_t *s_sum,
                                              const uint32_t *sum, uint32_t n) {
  const uint32_t nw = gsum_words(n, W);
  for (uint32_t i = threadIdx.x; i < nw; i += NT) s_sum[i] = sum[i];
  __syncthreads();
  return nw * 1024u;
}

// The LDS summary's answer for x: true = x is in the giant component; false
// leaves it open (the bitmap decides).
__device__ __forceinline__ bool giant_summary_hit(uint32_t x,
                                                  const uint32_t *s_sum,
                                                  uint32_t covered) {
  const uint32_t w = x >> 5;
  return x < covered && ((s_sum[w >> 5] >> (w & 31)) & 1u);
}

__device__ __forceinline__ bool giant_test(uint32_t x, const uint32_t *s_sum,
                                           uint32_t covered,
                                           const uint32_t *__restrict__ bits) {
  const uint32_t w = x >> 5;
  if (x < covered && ((s_sum[w >> 5] >> (w & 31)) & 1u)) return true;
  return (bits[w] >> (x & 31)) & 1u;
}

}  // namespace
}  // namespace phem

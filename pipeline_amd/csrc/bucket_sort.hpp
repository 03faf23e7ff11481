// LDS radix sort of one bucket of 32-bit keys by one workgroup, shared by
// the learned-CDF bucket sort (k_learned.hip) and the MSD sort of a build's
// keys (k_msd.hip).
#pragma once

#include "ocp_internal.hpp"

namespace ocp {

// Sorts g[0, n) in place, 0 < n <= WAVES * 64 * KPT: keys are loaded
// wave-contiguously, taken relative to the bucket minimum, and sorted by LSD
// passes over the bits they span - a counting sort with LDS atomics over the
// low 10-12 bits (one counter per digit in WAVES * 256 words), then stable
// 8-bit passes ranked per wave by ballot multisplit - scattered through LDS.
// Keys spanning 20 bits take two passes. Every thread of the workgroup
// calls it.
// With span > 0 the caller knows every key lies in [base, base + 2^span)
// and the workgroup skips finding the bucket's range.
template <int WAVES, int KPT, int P0 = (WAVES >= 16 ? 12 : WAVES == 8 ? 11 : 10)>
__device__ __forceinline__ void lds_bucket_sort(uint32_t* __restrict__ g, int32_t n,
                                                uint32_t base = 0, int span = 0) {
  constexpr int THREADS = WAVES * 64;
  constexpr int CAP = THREADS * KPT;
  // First-pass counters (2^P0) and the stable passes' per-wave digit
  // counters share one array.
  constexpr int NB = 1 << P0;
  constexpr int HW = NB > WAVES * 256 ? NB : WAVES * 256;
  static_assert(P0 >= 8 && NB % THREADS == 0, "whole counters per thread");
  __shared__ uint32_t s_keys[CAP];
  __shared__ uint32_t s_whist[HW];
  __shared__ uint32_t s_dstart[256];
  __shared__ uint32_t s_tmp[WAVES];
  __shared__ uint32_t s_mm[2];
  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  if (tid == 0) {
    s_mm[0] = 0xffffffffu;
    s_mm[1] = 0u;
  }
  uint32_t kv[KPT];
  uint32_t mn = 0xffffffffu, mx = 0u;
  // All loads first (consuming each in its own iteration keeps one load in
  // flight per wave), then the bucket's range.
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int32_t idx = static_cast<int32_t>(w) * (64 * KPT) + k * 64 + static_cast<int32_t>(lane);
    kv[k] = idx < n ? g[idx] : 0u;
  }
  int bits = span;
  if (span > 0) {
    mn = base;
  } else {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t idx = static_cast<int32_t>(w) * (64 * KPT) + k * 64 + static_cast<int32_t>(lane);
      if (idx < n) {
        mn = min(mn, kv[k]);
        mx = max(mx, kv[k]);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mn), o, 64)));
      mx = max(mx, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), o, 64)));
    }
    __syncthreads();
    if (lane == 0) {
      atomicMin(&s_mm[0], mn);
      atomicMax(&s_mm[1], mx);
    }
    __syncthreads();
    mn = s_mm[0];
    mx = s_mm[1];
    if (mn == mx) return;  // constant bucket: already sorted
    bits = 32 - __clz(static_cast<int>(mx - mn));
  }
  // Each wave owns KPT rows of 64 consecutive keys; a row that starts at or
  // past n holds padding only and is left out of every pass (its wave skips
  // the work), so a bucket far below CAP costs in proportion to its size.
  const int32_t wbase = static_cast<int32_t>(w) * (64 * KPT);
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int32_t idx = wbase + k * 64 + static_cast<int32_t>(lane);
    kv[k] = idx < n ? kv[k] - mn : 0xffffffffu;  // padding ranks last
  }
  const uint64_t lt = lanemask_lt();
  uint32_t* wh = s_whist + w * 256;
  // Pass 0 needs no stability (no earlier order to keep), so it is a plain
  // counting sort with LDS atomics over the low P0 bits, one counter per
  // word of s_whist; padding stays out of it, and the rest of the last
  // partial row is filled with it, where the stable passes keep it last.
  constexpr int PER = NB / THREADS;
  {
    for (int i = tid; i < NB; i += THREADS) s_whist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (wbase + k * 64 >= n) continue;
      const int32_t idx = wbase + k * 64 + static_cast<int32_t>(lane);
      if (idx < n) atomicAdd(&s_whist[kv[k] & (NB - 1u)], 1u);
    }
    __syncthreads();
    uint32_t cc[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      cc[j] = s_whist[tid * PER + j];
      sum += cc[j];
    }
    uint32_t all;
    uint32_t run = block_excl_scan<THREADS>(sum, s_tmp, &all);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      s_whist[tid * PER + j] = run;  // cursor per digit
      run += cc[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (wbase + k * 64 >= n) continue;
      const int32_t idx = wbase + k * 64 + static_cast<int32_t>(lane);
      if (idx < n) s_keys[atomicAdd(&s_whist[kv[k] & (NB - 1u)], 1u)] = kv[k];
    }
    const int32_t pad_end = (n + 63) & ~63;  // <= CAP
    for (int i = n + tid; i < pad_end; i += THREADS) s_keys[i] = 0xffffffffu;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k)
      if (wbase + k * 64 < n) kv[k] = s_keys[wbase + k * 64 + static_cast<int32_t>(lane)];
    __syncthreads();
  }
  for (int shift = P0; shift < bits; shift += 8) {
    for (int i = tid; i < WAVES * 256; i += THREADS) s_whist[i] = 0;
    __syncthreads();
    uint32_t off[KPT];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (wbase + k * 64 >= n) continue;
      const uint32_t d = (kv[k] >> shift) & 255u;
      uint64_t msk = ~0ull;
#pragma unroll
      for (int bb = 0; bb < 8; ++bb) {
        const bool bit = (d >> bb) & 1u;
        const uint64_t bal = __ballot(bit);
        msk &= bit ? bal : ~bal;
      }
      const uint32_t rank = __popcll(msk & lt);
      const uint32_t old = wh[d];
      if (rank == 0) wh[d] = old + __popcll(msk);
      off[k] = old + rank;
    }
    __syncthreads();
    uint32_t tot = 0;
    if (tid < 256) {
      uint32_t acc = 0;
#pragma unroll
      for (int ww = 0; ww < WAVES; ++ww) {
        const uint32_t cw = s_whist[ww * 256 + tid];
        s_whist[ww * 256 + tid] = acc;
        acc += cw;
      }
      tot = acc;
    }
    uint32_t all;
    const uint32_t ds = block_excl_scan<THREADS>(tot, s_tmp, &all);
    if (tid < 256) s_dstart[tid] = ds;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (wbase + k * 64 >= n) continue;
      const uint32_t d = (kv[k] >> shift) & 255u;
      s_keys[s_dstart[d] + wh[d] + off[k]] = kv[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; ++k)
      if (wbase + k * 64 < n) kv[k] = s_keys[wbase + k * 64 + static_cast<int32_t>(lane)];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int32_t idx = wbase + k * 64 + static_cast<int32_t>(lane);
    if (idx < n) g[idx] = kv[k] + mn;
  }
}

}  // namespace ocp

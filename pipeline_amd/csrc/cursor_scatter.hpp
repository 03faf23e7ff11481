// A scatter by digit whose order within a digit's run is free: the MSD
// sort's scatter by the top 10-bit digit (k_msd.hip) and the learned sort's
// two scatters by bucket (k_learned.hip), whose buckets are all sorted
// afterwards. Shared by both.
//
// Such a scatter needs no lookback. A tile (one workgroup) counts its keys
// per digit with LDS atomics, whose returned value is the key's rank in the
// tile's run of that digit; one thread per digit then claims the run's place
// in the segment with one atomic add on the digit's cursor (zeroed before
// the launch), on top of the digit's start in the segment, an exclusive scan
// of the segment's digit counts. Keys are staged in LDS in digit order and
// written with consecutive threads on consecutive slots of each run.
#pragma once

#include <type_traits>

#include "ocp_internal.hpp"

namespace ocp {

// The top 10-bit digit of a 30-bit key.
struct TopDigit10 {
  static constexpr bool kCheap = true;
  static constexpr int kBits = 10;
  __device__ __forceinline__ void setup(int32_t, int, int) const {}
  __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
    return (key >> 20) & 1023u;
  }
};

// Digit functor: kBits, kCheap (recomputed from the key rather than staged
// in LDS), setup(seg, tid, threads) for per-segment LDS tables, and the
// digit of a key. hist[seg * hstride + d]: the segment's count of digit d;
// cursor[seg * 2^kBits + d]: zeroed claim counters; tiles_dev (optional):
// the number of real tiles when the grid is sized for an upper bound. Tiles
// of THREADS * KPT keys never straddle a segment.
template <int THREADS, int KPT, class Digit>
__global__ __launch_bounds__(THREADS) void cursor_scatter_kernel(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, SegView sv,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tiles_dev,
    const uint32_t* __restrict__ hist, int64_t hstride, uint32_t* __restrict__ cursor,
    const uint32_t* __restrict__ ctrl, Digit dig) {
  constexpr int RADIX = 1 << Digit::kBits;
  constexpr int T = THREADS * KPT;
  constexpr bool CHEAP = Digit::kCheap;
  static_assert(THREADS >= RADIX, "one thread per digit");
  using DigT = typename std::conditional<(RADIX > 256), uint16_t, uint8_t>::type;
  __shared__ uint32_t s_keys[T];
  __shared__ DigT s_dig[CHEAP ? 1 : T];
  __shared__ uint32_t s_run[RADIX];  // tile counts, then the LDS start of each run
  __shared__ int32_t s_gpos[RADIX];  // global slot of a run's LDS slot 0
  __shared__ uint32_t s_tmp[2 * (THREADS / 64)];
  const int32_t t = static_cast<int32_t>(blockIdx.x);
  if (tiles_dev && t >= *tiles_dev) return;
  if (stopped_any(ctrl)) return;
  const int tid = threadIdx.x;
  const TileRef r = find_tile(tile_start, sv.n_seg, t);
  const int32_t sb = seg_begin(sv, r.seg);
  const int32_t len = seg_end(sv, r.seg) - sb;
  const int32_t base = sb + r.local * T;
  const int32_t cnt = min(T, sb + len - base);
  if (tid < RADIX) s_run[tid] = 0;
  dig.setup(r.seg, tid, THREADS);
  const uint32_t hv = tid < RADIX ? hist[static_cast<int64_t>(r.seg) * hstride + tid] : 0u;
  uint32_t keys[KPT];
  const uint32_t* src = in + base + tid;
  if (cnt == T) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) keys[k] = src[k * THREADS];
  } else {
#pragma unroll
    for (int k = 0; k < KPT; ++k) keys[k] = k * THREADS + tid < cnt ? src[k * THREADS] : 0u;
  }
  __syncthreads();
  uint32_t dg[KPT], rank[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    if (k * THREADS + tid < cnt) {
      dg[k] = dig(keys[k]);
      rank[k] = atomicAdd(&s_run[dg[k]], 1u);
    }
  }
  __syncthreads();
  const uint32_t c = tid < RADIX ? s_run[tid] : 0u;
  uint32_t lstart, gbase;
  block_excl_scan2<THREADS>(c, hv, s_tmp, &lstart, &gbase);
  if (tid < RADIX) {
    int32_t gpos = 0;
    if (c)
      gpos = sb +
             static_cast<int32_t>(gbase +
                                  atomicAdd(&cursor[static_cast<int64_t>(r.seg) * RADIX + tid], c)) -
             static_cast<int32_t>(lstart);
    s_run[tid] = lstart;  // (every read of the counts came before the scan's barriers)
    s_gpos[tid] = gpos;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    if (k * THREADS + tid < cnt) {
      const uint32_t pos = s_run[dg[k]] + rank[k];
      s_keys[pos] = keys[k];
      if constexpr (!CHEAP) s_dig[pos] = static_cast<DigT>(dg[k]);
    }
  }
  __syncthreads();
  for (int32_t i = tid; i < cnt; i += THREADS) {
    const uint32_t key = s_keys[i];
    const uint32_t d = CHEAP ? dig(key) : static_cast<uint32_t>(s_dig[i]);
    const int32_t p = s_gpos[d] + i;
    // (inside the segment whenever the counts describe these keys, as the
    // launchers' counts do)
    if (static_cast<uint32_t>(p - sb) < static_cast<uint32_t>(len)) out[p] = key;
  }
}

}  // namespace ocp

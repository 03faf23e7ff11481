// One stable OneSweep digit pass, generic in how a key maps to its 8-bit
// digit (a radix digit for the LSD sort, a learned-CDF bucket for the
// learned sort). Shared by k_onesweep.hip and k_learned.hip.
//
//   1. Tile = workgroup id. Workgroups are dispatched in id order on each
//      XCD, so the lowest unfinished tile is always resident and the
//      lookback makes progress without a shared ticket counter (a single
//      device-scope counter serialises at one atomic per tile).
//   2. Keys are loaded wave-contiguously (lane-strided, coalesced) and
//      ranked per wave with a 64-lane multisplit: one ballot per digit bit builds the
//      mask of lanes holding the same digit; popcount below the lane gives a
//      stable rank; the lowest lane bumps the wave's LDS digit counter.
//   3. The tile's 2^bits digit counts (256, or 1024 for 10-bit digits) are
//      published at once (aggregate, or inclusive for a segment's first
//      tile), then each of that many threads looks
//      back along its digit's chain, kOneSweepWindow tiles per round trip, with
//      agent-scope loads, and publishes the inclusive prefix.
//   4. Keys are scattered into LDS in digit order (with their digits, unless
//      a digit is cheap to recompute), then written out with consecutive
//      threads on consecutive LDS slots, so equal-digit runs go to
//      consecutive HBM addresses.
// Segments sort independently: each has its own digit histogram and its own
// lookback chain, and no tile holds keys of two segments.
//
// KV = true carries a 32-bit value with every key (the key->point
// permutation): values are loaded, staged in LDS and written beside their
// keys; with vin == nullptr a value is its key's index in the segment. Such
// passes are always stable, so that the values end as the stable argsort.
//
// STABLE = false is for passes whose order among equal digits is free: the
// first pass of an LSD sort of bare keys (its input order is arbitrary, so
// the sorted result is the same) and scatters whose buckets are sorted
// afterwards. Full tiles then rank each key with one LDS atomic add on the
// wave's digit counter instead of the eight-ballot multisplit; partial tiles
// keep the multisplit so that padding keys still rank last.
#pragma once

#include "ocp_internal.hpp"

namespace ocp {

#ifndef OCP_ONESWEEP_WINDOW
#define OCP_ONESWEEP_WINDOW 4
#endif
// Predecessor tiles polled per lookback round trip. Every poll of an
// unready window is wasted agent-scope traffic, so a short window measured
// faster than 8, 16, 32 or 64 (2^26 keys: 1.44 ms vs 1.52, 1.85, 2.2, 3.3).
constexpr int kOneSweepWindow = OCP_ONESWEEP_WINDOW;

// A digit functor: setup(seg, tid, threads) runs once per tile before any
// digit is taken (per-segment tables go to LDS there); operator() maps a key
// to its digit in [0, 2^kBits).
struct RadixDigit {
  // A digit this cheap is recomputed from the key in the scatter rather
  // than staged in LDS beside it.
  static constexpr bool kCheap = true;
  static constexpr int kBits = 8;
  uint32_t shift;
  __device__ __forceinline__ void setup(int32_t, int, int) const {}
  __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
    return (key >> shift) & 255u;
  }
};

// 10-bit digits: three passes cover a 30-bit Morton key.
struct RadixDigit10 {
  static constexpr bool kCheap = true;
  static constexpr int kBits = 10;
  uint32_t shift;
  __device__ __forceinline__ void setup(int32_t, int, int) const {}
  __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
    return (key >> shift) & 1023u;
  }
};

// hist: per segment hist_stride(kBits) words; this pass reads the 2^kBits
// words from slot * 2^kBits of its segment. tiles_dev (optional): the number
// of real tiles, on device, when the grid is sized for an upper bound.
template <int WAVES, int KPT, class Digit, bool STABLE = true, bool KV = false>
__global__ __launch_bounds__(WAVES * 64) void onesweep_kernel(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, SegView sv,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tiles_dev,
    const uint32_t* __restrict__ hist, int hist_slot, uint64_t* __restrict__ flags,
    uint32_t* __restrict__ ctrl, uint32_t stage, Digit dig,
    const uint32_t* __restrict__ vin = nullptr, uint32_t* __restrict__ vout = nullptr) {
  constexpr int THREADS = WAVES * 64;
  constexpr int T = THREADS * KPT;
  constexpr int BITS = Digit::kBits;
  constexpr int RADIX = 1 << BITS;
  constexpr int64_t HSTRIDE = hist_stride(BITS);
  static_assert(THREADS >= RADIX, "one thread per digit is required");
  static_assert(!KV || STABLE, "a pass that carries values is stable");
  constexpr bool CHEAP = Digit::kCheap;
  static_assert(CHEAP || BITS == 8, "staged digits are bytes");
  __shared__ uint32_t s_keys[T];
  __shared__ uint32_t s_vals[KV ? T : 1];
  __shared__ uint8_t s_dig[CHEAP ? 1 : T];
  __shared__ uint32_t s_whist[WAVES * RADIX];
  __shared__ int32_t s_dst[RADIX];
  __shared__ uint32_t s_tmp[2 * WAVES];
  __shared__ uint32_t s_trivial;

  const int32_t t = blockIdx.x;
  if (tiles_dev && t >= *tiles_dev) return;
  const uint32_t tag = ctrl[kCtrlEpoch] * 8u + stage;
  if (stopped(ctrl, tag)) return;
  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  const TileRef r = find_tile(tile_start, sv.n_seg, t);
  for (int i = tid; i < WAVES * RADIX; i += THREADS) s_whist[i] = 0;
  if (tid == 0) s_trivial = 0u;
  dig.setup(r.seg, tid, THREADS);
  __syncthreads();
  const int32_t sb = seg_begin(sv, r.seg);
  const int32_t base = sb + r.local * T;
  const int32_t cnt = min(T, seg_end(sv, r.seg) - base);
  // The segment's count of this thread's digit, loaded now so that it is in
  // hand when the tile's offsets are formed.
  const uint32_t hv =
      tid < RADIX ? hist[static_cast<int64_t>(r.seg) * HSTRIDE + hist_slot * RADIX + tid] : 0u;

  uint32_t keys[KPT];
  uint32_t dg[KPT];
  uint32_t off[KPT];
  uint32_t vals[KV ? KPT : 1];
  const int32_t lbase = static_cast<int32_t>(w) * (64 * KPT) + static_cast<int32_t>(lane);
  const uint32_t* src = in + lbase + base;
  if constexpr (KV) {
    // Values: given, or the key's index in its segment (the first pass of a
    // build, where the keys are in point order).
    const uint32_t* vsrc = vin ? vin + lbase + base : nullptr;
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t idx = lbase + k * 64;
      vals[k] = vsrc ? (idx < cnt ? vsrc[k * 64] : 0u)
                     : static_cast<uint32_t>(base - sb + idx);
    }
  }
  if (cnt == T) {  // every tile but a segment's last: no per-key guards
#pragma unroll
    for (int k = 0; k < KPT; ++k) keys[k] = src[k * 64];
#pragma unroll
    for (int k = 0; k < KPT; ++k) dg[k] = dig(keys[k]);
  } else {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t idx = static_cast<int32_t>(w) * (64 * KPT) + k * 64 + static_cast<int32_t>(lane);
      keys[k] = idx < cnt ? src[k * 64] : 0xffffffffu;
    }
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t idx = static_cast<int32_t>(w) * (64 * KPT) + k * 64 + static_cast<int32_t>(lane);
      dg[k] = idx < cnt ? dig(keys[k]) : uint32_t(RADIX - 1);  // padding keys rank last
    }
  }

  // A segment whose keys all share this pass's digit (one bin of its
  // histogram holds every key) needs no pass: a stable pass over it is the
  // identity. The thread of that bin flags it; after the barrier that
  // follows the ranking, the tile copies its keys through. Every tile of the
  // segment decides the same way, so none waits on a flag never published.
  if (tid < RADIX && static_cast<int32_t>(hv) == seg_end(sv, r.seg) - sb) s_trivial = 1u;

  // Wave-level multisplit. The read of the digit counter by every lane of a
  // group happens in the same LDS instruction, before the leader's update.
  uint32_t* wh = s_whist + w * RADIX;
  if (STABLE || cnt < T) {
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const uint32_t d = dg[k];
      uint64_t m = ~0ull;
#pragma unroll
      for (int b = 0; b < BITS; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
      }
      const uint32_t rank = __popcll(m & lt);
      const uint32_t old = wh[d];
      if (rank == 0) wh[d] = old + __popcll(m);
      off[k] = old + rank;
    }
  } else {
    // Each lane takes the next free rank of its digit within the wave.
    lds_ptr<uint32_t> whl = as_lds<uint32_t>(wh);
#pragma unroll
    for (int k = 0; k < KPT; ++k)
      off[k] = __hip_atomic_fetch_add(whl + dg[k], 1u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  if (s_trivial) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t idx = lbase + k * 64;
      if (idx < cnt) {
        out[base + idx] = keys[k];
        if constexpr (KV) vout[base + idx] = vals[k];
      }
    }
    return;
  }

  // Per-digit: exclusive prefix over waves, tile count, publish.
  uint32_t tile_cnt = 0;
  if (tid < RADIX) {
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) {
      const uint32_t c = s_whist[ww * RADIX + tid];
      s_whist[ww * RADIX + tid] = run;
      run += c;
    }
    if (tid == RADIX - 1) run -= static_cast<uint32_t>(T - cnt);  // padding keys
    tile_cnt = run;
    flag_store(&flags[static_cast<int64_t>(t) * RADIX + tid],
               flag_pack(tag, r.local == 0 ? kFlagInc : kFlagAgg, run));
  }
  // Tile-local digit starts and the segment's digit starts.
  uint32_t dstart, hex;
  block_excl_scan2<THREADS>(tile_cnt, hv, s_tmp, &dstart, &hex);
  if (tid < RADIX) {
    // Each wave's offset of this digit becomes its LDS position in the
    // tile: digit start + keys of the digit in earlier waves.
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) s_whist[ww * RADIX + tid] += dstart;
    uint32_t prev = 0;
    if (r.local > 0) {
      bool late = false;
      prev = lookback_n<kOneSweepWindow>(flags + tid, RADIX, t, t - r.local, tag,
                                         wait_limit(ctrl), &late);
      if (late) raise_timeout(ctrl, tag);
      flag_store(&flags[static_cast<int64_t>(t) * RADIX + tid],
                 flag_pack(tag, kFlagInc, prev + tile_cnt));
    }
    s_dst[tid] = sb + static_cast<int32_t>(hex + prev) - static_cast<int32_t>(dstart);
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const uint32_t pos = wh[dg[k]] + off[k];
    s_keys[pos] = keys[k];
    if constexpr (KV) s_vals[pos] = vals[k];
    if constexpr (!CHEAP) s_dig[pos] = static_cast<uint8_t>(dg[k]);
  }
  __syncthreads();

  if (cnt == T) {
    // All LDS reads first, then the stores (a guarded store per iteration
    // waits on its two dependent LDS reads before the next begins).
    uint32_t kv[KPT];
    uint32_t vv[KV ? KPT : 1];
    int32_t dst[KPT];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      kv[k] = s_keys[k * THREADS + tid];
      if constexpr (KV) vv[k] = s_vals[k * THREADS + tid];
      if constexpr (!CHEAP) dst[k] = s_dig[k * THREADS + tid];
    }
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if constexpr (CHEAP) dst[k] = static_cast<int32_t>(dig(kv[k]));
      dst[k] = s_dst[dst[k]] + k * THREADS + tid;
    }
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      out[dst[k]] = kv[k];
      if constexpr (KV) vout[dst[k]] = vv[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int32_t p = k * THREADS + tid;
      if (p < cnt) {
        const uint32_t kv = s_keys[p];
        const uint32_t d = CHEAP ? dig(kv) : s_dig[p];
        out[s_dst[d] + p] = kv;
        if constexpr (KV) vout[s_dst[d] + p] = s_vals[p];
      }
    }
  }
}

}  // namespace ocp

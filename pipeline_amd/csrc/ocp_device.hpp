// Device-side building blocks shared by every stage kernel (gfx950, wave64).
//
// - morton codec: the reference's 10-bit-per-axis encode and decode
//   (include/shared/morton.h:15-80), integer-exact, float parts compiled with
//   -ffp-contract=off so products and sums round separately;
// - wave64 scans on ds_swizzle/bpermute, block scans over LDS;
// - decoupled-lookback flags: 64-bit words {tag:30, status:2, count:32}
//   published and polled with agent-scope atomics so that they are coherent
//   across the eight XCD-private L2s; the tag carries an epoch so that stale
//   flags from an earlier launch are never mistaken for ready ones and the
//   flag array never needs clearing;
// - the segment view: every stage kernel works on a batch of independent
//   clouds described by a point-offset table (one cloud = one segment); a
//   tile never straddles two segments.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocp {

constexpr int kWave = 64;
constexpr int kMortonBits = 30;

// ---------------------------------------------------------------- LDS pointers
// Pointers typed as LDS (address space 3). A window accessor such as
// `in_window ? lds[q] : global[q]` through generic pointers lets the
// compiler fold both loads into one FLAT load, which goes through the vector
// memory path even for LDS; with the LDS pointer typed, the window hit is a
// ds_read.
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;
template <class T>
__device__ __forceinline__ lds_ptr<T> as_lds(T* p) {
  return (lds_ptr<T>)(p);
}

// ---------------------------------------------------------------- lanes
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x >> 6; }

// ---------------------------------------------------------------- morton
__device__ __forceinline__ uint32_t split_by_3(uint32_t a) {
  uint32_t x = a & 0x3ffu;
  x = (x | x << 16) & 0x30000ffu;
  x = (x | x << 8) & 0x0300f00fu;
  x = (x | x << 4) & 0x30c30c3u;
  x = (x | x << 2) & 0x9249249u;
  return x;
}
__device__ __forceinline__ uint32_t third_bits(uint32_t m) {
  uint32_t x = m & 0x9249249u;
  x = (x ^ (x >> 2)) & 0x30c30c3u;
  x = (x ^ (x >> 4)) & 0x0300f00fu;
  x = (x ^ (x >> 8)) & 0x30000ffu;
  x = (x ^ (x >> 16)) & 0x3ffu;
  return x;
}
// (a - min) / range * 1024, truncated toward zero: v_sub, the correctly
// rounded v_div sequence, v_mul, v_cvt_i32. Going through int32 gives the
// same bits as the x86 truncating conversion for coordinates a little below
// min as well (the reference leaves those undefined).
__device__ __forceinline__ uint32_t quantize(float a, float mn, float rg) {
  const float v = (a - mn) / rg * 1024.0f;
  return static_cast<uint32_t>(static_cast<int32_t>(v));
}
__device__ __forceinline__ uint32_t encode(float x, float y, float z, float mn,
                                           float rg) {
  return split_by_3(quantize(x, mn, rg)) |
         (split_by_3(quantize(y, mn, rg)) << 1) |
         (split_by_3(quantize(z, mn, rg)) << 2);
}
__device__ __forceinline__ float decode_axis(uint32_t code, int axis, float mn,
                                             float rg) {
  const float t = static_cast<float>(third_bits(code >> axis)) / 1024.0f;
  return t * rg + mn;
}

// ---------------------------------------------------------------- scans
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, kWave);
    if (l >= static_cast<uint32_t>(o)) v += t;
  }
  return v;
}
__device__ __forceinline__ int32_t wave_incl_max(int32_t v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

// Exclusive scan of one value per thread over a block of THREADS threads.
// `tmp` holds THREADS/64 words of LDS. Ends with a barrier so `tmp` can be
// reused right away.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tmp,
                                                    uint32_t* total) {
  constexpr int NW = THREADS / kWave;
  const uint32_t incl = wave_incl_scan(v);
  const uint32_t w = wave_id();
  if (lane_id() == kWave - 1) tmp[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const uint32_t t = tmp[k];
    off += (static_cast<uint32_t>(k) < w) ? t : 0u;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Two exclusive scans at once (one barrier pair): `tmp` holds 2 * THREADS/64
// words of LDS. Ends with a barrier so `tmp` can be reused right away.
template <int THREADS>
__device__ __forceinline__ void block_excl_scan2(uint32_t a, uint32_t b, uint32_t* tmp,
                                                 uint32_t* ea, uint32_t* eb) {
  constexpr int NW = THREADS / kWave;
  const uint32_t ia = wave_incl_scan(a);
  const uint32_t ib = wave_incl_scan(b);
  const uint32_t w = wave_id();
  if (lane_id() == kWave - 1) {
    tmp[w] = ia;
    tmp[NW + w] = ib;
  }
  __syncthreads();
  uint32_t oa = 0, ob = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const uint32_t ta = tmp[k];
    const uint32_t tb = tmp[NW + k];
    oa += (static_cast<uint32_t>(k) < w) ? ta : 0u;
    ob += (static_cast<uint32_t>(k) < w) ? tb : 0u;
  }
  __syncthreads();
  *ea = oa + ia - a;
  *eb = ob + ib - b;
}

// Exclusive scan, in place, of R words of LDS by one whole wave (each lane
// takes R/64 consecutive words); returns the total in every lane.
template <int R>
__device__ __forceinline__ uint32_t wave_scan_lds(uint32_t* a) {
  constexpr int PER = (R + kWave - 1) / kWave;
  const uint32_t lane = lane_id();
  uint32_t v[PER];
  uint32_t sum = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const uint32_t idx = lane * PER + q;
    v[q] = idx < static_cast<uint32_t>(R) ? a[idx] : 0u;
    sum += v[q];
  }
  const uint32_t incl = wave_incl_scan(sum);
  uint32_t run = incl - sum;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const uint32_t idx = lane * PER + q;
    if (idx < static_cast<uint32_t>(R)) a[idx] = run;
    run += v[q];
  }
  return __shfl(incl, kWave - 1, kWave);
}

// ---------------------------------------------------------------- lookback
constexpr uint32_t kFlagAgg = 1;  // the tile's own count is published
constexpr uint32_t kFlagInc = 2;  // the inclusive prefix is published

__device__ __forceinline__ uint64_t flag_pack(uint32_t tag, uint32_t status,
                                              uint32_t count) {
  return (static_cast<uint64_t>(((tag & 0x3fffffffu) << 2) | status) << 32) |
         count;
}
__device__ __forceinline__ void flag_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t flag_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
// Every wait below is bounded: after `limit` unsuccessful polls (kSpinLimit
// unless the host sets another) the caller gets `timed_out` set, publishes
// what it has and the kernel runs to its end; the host turns the flag into
// an error. Tiles are numbered by workgroup id, and workgroups are
// dispatched in id order on each XCD, so the lowest unfinished tile is
// always resident and a wait ends long before the limit.
constexpr uint32_t kSpinLimit = 1u << 22;

// Lookback of one thread per chain (one digit of a OneSweep tile), polling
// W predecessors per round trip: flags of tiles k-1..k-W are loaded together
// and consumed in order while ready. `first` is the chain's first tile,
// which always publishes an inclusive prefix.
template <int W>
__device__ __forceinline__ uint32_t lookback_n(const uint64_t* flags,
                                               uint32_t stride, int32_t t,
                                               int32_t first, uint32_t tag,
                                               uint32_t limit, bool* timed_out) {
  uint32_t excl = 0;
  int32_t k = t - 1;
  const uint32_t want = tag & 0x3fffffffu;
  for (uint32_t spin = 0; spin < limit;) {
    uint64_t v[W];
#pragma unroll
    for (int u = 0; u < W; ++u)
      v[u] = (k - u >= first) ? flag_load(flags + static_cast<int64_t>(k - u) * stride) : 0ull;
    int consumed = 0;
    bool stop = false;
#pragma unroll
    for (int u = 0; u < W; ++u) {
      if (stop) continue;
      const uint32_t hi = static_cast<uint32_t>(v[u] >> 32);
      const uint32_t st = hi & 3u;
      if ((hi >> 2) != want || st == 0) {
        stop = true;
        continue;
      }
      excl += static_cast<uint32_t>(v[u]);
      if (st == kFlagInc) return excl;
      ++consumed;
    }
    if (consumed == 0) {
      __builtin_amdgcn_s_sleep(1);
      ++spin;
    }
    k -= consumed;
  }
  *timed_out = true;
  return excl;
}

// Sum (or, with MAX, maximum) of the counts of `cnt` (< 65) flags base[0],
// base[stride], ... by a whole wave once every one of them carries `want`;
// bounded like every wait.
template <bool MAX = false>
__device__ __forceinline__ uint32_t wave_sum_flags(const uint64_t* base, int32_t stride,
                                                   int32_t cnt, uint32_t want,
                                                   uint32_t limit, bool* timed_out) {
  const int32_t lane = static_cast<int32_t>(lane_id());
  for (uint32_t spin = 0; spin < limit; ++spin) {
    bool ready = true;
    uint32_t c = 0;
    if (lane < cnt) {
      const uint64_t v = flag_load(base + static_cast<int64_t>(lane) * stride);
      const uint32_t hi = static_cast<uint32_t>(v >> 32);
      ready = (hi >> 2) == want && (hi & 3u) != 0;
      c = static_cast<uint32_t>(v);
    }
    if (__ballot(!ready)) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t x = __shfl_xor(c, o, kWave);
      c = MAX ? max(c, x) : c + x;
    }
    return c;
  }
  *timed_out = true;
  return 0;
}

// Two-level lookback of a single-value chain by a whole wave. Every tile
// publishes its own count at flags[t] before calling; tiles are grouped by
// 64 from the chain's first tile, and the last tile of a group also
// publishes the group's sum at gflags[first tile of the group]. Tile t sums
// the counts of its group's earlier tiles and the sums of all earlier
// groups: two round trips (plus one per 64 groups) whatever its position,
// where a chain of single-tile flags resolves about 128 tiles per round
// trip. No wait depends on a later tile: the counts are published before
// any wait, and a group's last tile waits only on its own group. Every lane
// must call it; all lanes return the exclusive prefix of tile t.
__device__ __forceinline__ uint32_t wave_lookback_groups(uint64_t* flags, uint64_t* gflags,
                                                         int32_t t, int32_t first,
                                                         uint32_t own, uint32_t tag,
                                                         uint32_t limit, bool* timed_out) {
  const uint32_t want = tag & 0x3fffffffu;
  const int32_t local = t - first;
  const int32_t g0 = local & ~63;  // this tile's group, as a local tile index
  uint32_t excl = wave_sum_flags(flags + first + g0, 1, local - g0, want, limit, timed_out);
  if ((local & 63) == 63 && lane_id() == 0)
    flag_store(gflags + first + g0, flag_pack(tag, kFlagAgg, excl + own));
  for (int32_t g = 0; g < g0; g += 64 * 64)
    excl += wave_sum_flags(gflags + first + g, 64, min(64, (g0 - g) >> 6), want, limit,
                           timed_out);
  return excl;
}


// ---------------------------------------------------------------- segments
// A batch of independent clouds. `start` holds n_seg+1 point offsets; with
// start == nullptr the batch is the single segment [0, n1) with box
// (min1, range1). Per-segment boxes come from mn/rg when non-null. With
// `end` non-null, segment k is [start[k], end[k]) and segments need not be
// adjacent (the learned sort's oversized buckets).
struct SegView {
  int32_t n_seg;
  const int32_t* start;
  const float* mn;
  const float* rg;
  int32_t n1;
  float min1;
  float range1;
  const int32_t* end;
};

__device__ __forceinline__ int32_t seg_begin(const SegView& s, int32_t k) {
  return s.start ? s.start[k] : 0;
}
__device__ __forceinline__ int32_t seg_end(const SegView& s, int32_t k) {
  return s.end ? s.end[k] : (s.start ? s.start[k + 1] : s.n1);
}
__device__ __forceinline__ float seg_min(const SegView& s, int32_t k) {
  return s.mn ? s.mn[k] : s.min1;
}
__device__ __forceinline__ float seg_range(const SegView& s, int32_t k) {
  return s.rg ? s.rg[k] : s.range1;
}

// Tile map for one tile size: tile_start[k] = first global tile of segment
// k (n_seg+1 entries, on device). nullptr means a single segment.
struct TileRef {
  int32_t seg;
  int32_t local;  // tile index inside its segment
};
__device__ __forceinline__ TileRef find_tile(const int32_t* tile_start,
                                             int32_t n_seg, int32_t t) {
  if (!tile_start) return {0, t};
  int32_t lo = 0, hi = n_seg;  // invariant: tile_start[lo] <= t < tile_start[hi]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (tile_start[mid] <= t) lo = mid; else hi = mid;
  }
  return {lo, t - tile_start[lo]};
}

// Workgroups are dispatched round-robin over the chip's 8 XCDs (block b to
// XCD b % 8), and each XCD has its own L2. For kernels without cross-tile
// waits, this maps block b to a tile so that every XCD walks one contiguous
// run of tiles in order: the halo a tile shares with its neighbour is then
// still in that XCD's L2. Bijective on [0, g).
__device__ __forceinline__ int32_t xcd_tile(int32_t b, int32_t g) {
  const int32_t x = b & 7, k = b >> 3;
  const int32_t per = g >> 3, rem = g & 7;
  return x * per + min(x, rem) + k;
}

// Reference OctNode layout (include/shared/types.h:28-34), 60 bytes.
struct OctNode {
  int32_t children[8];
  float corner[4];
  float cell_size;
  int32_t child_node_mask;
  int32_t child_leaf_mask;
};
static_assert(sizeof(OctNode) == 60, "OctNode must stay 60 bytes");

}  // namespace ocp

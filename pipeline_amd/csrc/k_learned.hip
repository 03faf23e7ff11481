// Learned-CDF bucket sort of 32-bit keys, and the kernels that train and
// compile its model.
//
// Model (FP32, layout in learned.hpp): a two-level piecewise-linear CDF.
//   root  r(x) = b2 + sum_j w2_j * relu(w1_j * x + b1_j), 16 hidden units,
//         x = (key - kmin) / (kmax - kmin + 1); after every SGD step
//         w1_j * w2_j >= 0 is restored, so each term, and r, is
//         non-decreasing in x (float rounding keeps that: every operation
//         is monotone in its operands);
//   leaf  l(key) = clamp(floor(r * 4096), 0, 4095); leaf l starts at key
//         K_l = min{key : l(key) >= l} and maps
//         y = clamp(a_l + b_l * (key - K_l), lo_l, lo_{l+1}), b_l >= 0,
//         lo_l = the sample CDF at K_l;
//   bucket q(key) = clamp(floor(y * 16384), 0, 16383).
// q is non-decreasing in the key, so it is fully described by its 16384
// bucket boundaries fb[f] = min{key : q(key) >= f}: q(key) = #{f >= 1 :
// fb[f] <= key}. The compile step finds every fb[f] by binary search over
// the 32-bit key space with the model itself, and a table of the bucket at
// each of 8192 equal cells of the key range; the sort evaluates q only
// through those (one LDS read for most keys, a short search in LDS for keys
// in cells that hold a boundary), never the network.
//
// Sort (keys only, so stability is irrelevant and any correct per-bucket
// sort gives the bytes of np.sort):
//   1. histogram of q (persistent workgroups, fb and counters in LDS);
//   2. one workgroup turns it into bucket starts, the two scatter passes'
//      digit tables and the list of oversized buckets;
//   3. scatter by the coarse digit q >> 6, then per coarse segment by the
//      fine digit q & 63 (cursor_scatter.hpp: the order within a bucket is
//      free, so tiles claim their runs with atomic adds on per-digit
//      cursors instead of a lookback), which leaves every bucket contiguous
//      at its start;
//   4. one workgroup per bucket sorts it in LDS (LSD radix over the bits its
//      keys span above the bucket minimum), in four size classes;
//   5. buckets above 16384 keys are sorted by the segmented OneSweep
//      (host side); constant buckets need nothing;
//   6. every bucket boundary is checked (a misorder makes the host re-sort
//      the whole array with OneSweep).
#include "bucket_sort.hpp"
#include "cursor_scatter.hpp"
#include "learned.hpp"
#include "onesweep_impl.hpp"

namespace ocp {
namespace {

// ------------------------------------------------------------ model eval
struct CdfDev {
  const float* p;
  const uint32_t* kleaf;
  uint32_t kmin;
  float inv_span;
};

__device__ __forceinline__ CdfDev cdf_view(const float* p, const uint32_t* meta,
                                           const uint32_t* kleaf) {
  CdfDev v;
  v.p = p;
  v.kleaf = kleaf;
  v.kmin = meta[0];
  v.inv_span = 1.0f / (static_cast<float>(meta[1] - meta[0]) + 1.0f);
  return v;
}

__device__ __forceinline__ float cdf_x(const CdfDev& v, uint32_t key) {
  const uint32_t d = key > v.kmin ? key - v.kmin : 0u;
  return static_cast<float>(d) * v.inv_span;
}

__device__ __forceinline__ float root_eval(const float* __restrict__ m, float x) {
  float r = m[kCdfB2];
#pragma unroll
  for (int j = 0; j < kCdfHidden; ++j) {
    const float h = fmaxf(m[kCdfW1 + j] * x + m[kCdfB1 + j], 0.0f);
    r = r + m[kCdfW2 + j] * h;
  }
  return r;
}

// The maximal key always takes the last leaf and bucket, so every boundary
// exists inside the 32-bit key space.
__device__ __forceinline__ int32_t leaf_of(const CdfDev& v, uint32_t key) {
  if (key == 0xffffffffu) return kCdfLeaves - 1;
  const float s = floorf(root_eval(v.p, cdf_x(v, key)) * static_cast<float>(kCdfLeaves));
  if (!(s > 0.0f)) return 0;  // also NaN
  return s >= static_cast<float>(kCdfLeaves) ? kCdfLeaves - 1 : static_cast<int32_t>(s);
}

__device__ __forceinline__ uint32_t q_model(const CdfDev& v, uint32_t key) {
  if (key == 0xffffffffu) return kLearnBuckets - 1;
  const int32_t l = leaf_of(v, key);
  const float d = static_cast<float>(key - v.kleaf[l]);
  const float lo = v.p[kCdfLeafLo + l];
  const float hi = v.p[kCdfLeafLo + l + 1];
  float y = v.p[kCdfLeafA + l] + v.p[kCdfLeafB + l] * d;
  y = fminf(fmaxf(y, lo), hi);
  const float s = floorf(y * static_cast<float>(kLearnBuckets));
  if (!(s > 0.0f)) return 0u;
  return s >= static_cast<float>(kLearnBuckets) ? kLearnBuckets - 1u : static_cast<uint32_t>(s);
}

// Bounds [lo, hi] on q(key) from the cell table (q is non-decreasing, so a
// key's bucket lies between those of its cell's first key and the next
// cell's); keys outside the table's range get the open-ended bounds.
template <class P16>
__device__ __forceinline__ void cell_bounds(P16 ctab, uint32_t kmin, uint32_t cshift,
                                            uint32_t key, uint32_t& lo, uint32_t& hi) {
  if (key < kmin) {
    lo = 0;
    hi = ctab[0];
    return;
  }
  const uint32_t c = (key - kmin) >> cshift;
  if (c >= static_cast<uint32_t>(kLearnCells)) {
    lo = ctab[kLearnCells];
    hi = kLearnBuckets - 1;
    return;
  }
  lo = ctab[c];
  hi = ctab[c + 1];
}
// Largest f in [lo, hi] with tab[f - base] <= key, tab[lo - base] <= key.
template <class P32>
__device__ __forceinline__ uint32_t search_range(P32 tab, uint32_t base, uint32_t lo,
                                                 uint32_t hi, uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (tab[mid - base] <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Digit functors of the two scatter passes; their tables live in LDS.
__shared__ uint16_t s_ctab[kLearnCells + 2];
__shared__ uint32_t s_cb[kLearnCoarse];
__shared__ uint32_t s_fbs[kLearnFine];
__shared__ uint32_t s_kmin, s_cshift;

__device__ __forceinline__ void load_cells(const LearnTables& tb, int tid, int threads) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(tb.ctab);
  uint32_t* dst = reinterpret_cast<uint32_t*>(s_ctab);
  for (int i = tid; i < (kLearnCells + 2) / 2; i += threads) dst[i] = src[i];
  if (tid == 0) {
    s_kmin = tb.meta[0];
    s_cshift = tb.meta[2];
  }
}

// Per-key branch-free binary search over the 256 coarse boundaries (pass A)
// or the 64 boundaries of the tile's coarse segment (pass B). In the
// scatter passes this measured faster than the cell table, which every tile
// would have to load.
struct CoarseDigit {
  static constexpr bool kCheap = false;
  static constexpr int kBits = 8;
  LearnTables tb;
  __device__ __forceinline__ void setup(int32_t, int tid, int threads) const {
    for (int i = tid; i < kLearnCoarse; i += threads) s_cb[i] = tb.fb[i * kLearnFine];
  }
  __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
    const auto cb = as_lds<const uint32_t>(s_cb);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t step = kLearnCoarse / 2; step > 0; step >>= 1) c += cb[c + step] <= key ? step : 0u;
    return c;
  }
};
struct FineDigit {
  static constexpr bool kCheap = false;
  static constexpr int kBits = 8;
  LearnTables tb;
  __device__ __forceinline__ void setup(int32_t seg, int tid, int threads) const {
    for (int i = tid; i < kLearnFine; i += threads) s_fbs[i] = tb.fb[seg * kLearnFine + i];
  }
  __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
    // every key of the tile is in this coarse segment: fbs[0] <= key
    const auto fbs = as_lds<const uint32_t>(s_fbs);
    uint32_t f = 0;
#pragma unroll
    for (uint32_t step = kLearnFine / 2; step > 0; step >>= 1) f += fbs[f + step] <= key ? step : 0u;
    return f;
  }
};

// ------------------------------------------------------------ histogram
constexpr int kHistThreads = 1024;

__global__ __launch_bounds__(kHistThreads) void lhist_kernel(const uint32_t* __restrict__ keys,
                                                             int64_t n, LearnTables tb,
                                                             uint32_t* __restrict__ rows) {
  __shared__ uint32_t s_fb[kLearnBuckets];
  __shared__ uint32_t s_cnt[kLearnBuckets];
  const int tid = threadIdx.x;
  for (int i = tid; i < kLearnBuckets; i += kHistThreads) {
    s_fb[i] = tb.fb[i];
    s_cnt[i] = 0;
  }
  load_cells(tb, tid, kHistThreads);
  __syncthreads();
  const auto ctab = as_lds<const uint16_t>(s_ctab);
  const auto fb = as_lds<const uint32_t>(s_fb);
  const uint32_t kmin = s_kmin, cshift = s_cshift;
  auto count = [&](uint32_t key) {
    uint32_t lo, hi;
    cell_bounds(ctab, kmin, cshift, key, lo, hi);
    atomicAdd(&s_cnt[search_range(fb, 0u, lo, hi, key)], 1u);
  };
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kHistThreads;
  int64_t i = static_cast<int64_t>(blockIdx.x) * kHistThreads + tid;
  constexpr int U = 16;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint32_t k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = keys[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) count(k[u]);
  }
  for (; i < n; i += stride) count(keys[i]);
  __syncthreads();
  uint32_t* row = rows + static_cast<int64_t>(blockIdx.x) * kLearnBuckets;
  for (int q = tid; q < kLearnBuckets; q += kHistThreads) row[q] = s_cnt[q];
}

// Column sums of the matrix in kLearnReduceGroups groups of rows (enough
// workgroups to keep many loads in flight), then the group partials are
// added into the first row.
__global__ __launch_bounds__(256) void lreduce_kernel(const uint32_t* __restrict__ rows,
                                                      int32_t n_rows,
                                                      uint32_t* __restrict__ partial) {
  const int32_t q = blockIdx.x * 256 + threadIdx.x;
  const int32_t per = (n_rows + kLearnReduceGroups - 1) / kLearnReduceGroups;
  const int32_t r0 = blockIdx.y * per;
  const int32_t r1 = min(n_rows, r0 + per);
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t r = r0;
  for (; r + 8 <= r1; r += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += rows[static_cast<int64_t>(r + u) * kLearnBuckets + q];
  }
  for (; r < r1; ++r) acc[0] += rows[static_cast<int64_t>(r) * kLearnBuckets + q];
  partial[static_cast<int64_t>(blockIdx.y) * kLearnBuckets + q] =
      acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5] + acc[6] + acc[7];
}

__global__ __launch_bounds__(256) void lreduce2_kernel(uint32_t* __restrict__ hist) {
  const int32_t q = blockIdx.x * 256 + threadIdx.x;
  uint32_t v = 0;
#pragma unroll
  for (int gi = 0; gi < kLearnReduceGroups; ++gi) v += hist[gi * kLearnBuckets + q];
  hist[q] = v;
}

// ------------------------------------------------------------ tables
// One workgroup of 1024 threads, 16 buckets per thread.
__global__ __launch_bounds__(1024) void lprefix_kernel(LearnTables tb, int32_t fallback_tile) {
  __shared__ uint32_t tmp[32];
  __shared__ int32_t s_bs[kLearnBuckets + 1];
  __shared__ int32_t s_max;
  __shared__ uint32_t s_vary;
  constexpr int PER = kLearnBuckets / 1024;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_max = 0;
    s_vary = 0;
  }
  uint32_t v[PER];
  uint32_t sum = 0, nonempty = 0, big_keys = 0;
  // Keys land in bucket b exactly when fb[b] <= key < fb[b+1] (bucket 0
  // takes every key below fb[1], the last one every key from fb[16383] on:
  // the search returns the largest f with fb[f] <= key), so the bits above
  // the highest bit of lo ^ (hi - 1) are the same for every key of the
  // bucket and the fallback's passes over those digits are identities.
  // Oversized buckets whose range varies in the low two digits only form
  // group A (two passes), the others group B (four passes).
  uint32_t vary = 0, wide = 0;  // wide: bit q set for a group-B bucket
  uint32_t nbig_a = 0, nbig_b = 0, tiles_a = 0, tiles_b = 0;
  int32_t mx = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    v[q] = tb.hist[tid * PER + q];
    sum += v[q];
    nonempty += v[q] ? 1u : 0u;
    mx = max(mx, static_cast<int32_t>(v[q]));
    if (v[q] > static_cast<uint32_t>(kLearnBigCap)) {
      const int32_t b = tid * PER + q;
      const uint32_t lo = b == 0 ? 0u : tb.fb[b];
      const uint32_t hi = b + 1 < kLearnBuckets ? tb.fb[b + 1] : 0u;
      const uint32_t span =
          (b + 1 < kLearnBuckets && hi > lo) ? (lo ^ (hi - 1u)) : 0xffffffffu;
      vary |= span;
      const uint32_t tiles = (v[q] + fallback_tile - 1) / fallback_tile;
      big_keys += v[q];
      if (span >> 16) {
        wide |= 1u << q;
        ++nbig_b;
        tiles_b += tiles;
      } else {
        ++nbig_a;
        tiles_a += tiles;
      }
    }
  }
  uint32_t total;
  uint32_t run = block_excl_scan<1024>(sum, tmp, &total);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    s_bs[tid * PER + q] = static_cast<int32_t>(run);
    tb.bstart[tid * PER + q] = static_cast<int32_t>(run);
    run += v[q];
  }
  if (tid == 0) {
    s_bs[kLearnBuckets] = static_cast<int32_t>(total);
    tb.bstart[kLearnBuckets] = static_cast<int32_t>(total);
  }
  // (after the scan's barriers: s_max and s_vary are cleared before them)
  atomicMax(&s_max, mx);
  if (vary) atomicOr(&s_vary, vary);
  // Oversized buckets, in bucket order within each group: group A takes
  // segments [0, nA) and tile map big_tiles[0 .. nA], group B segments
  // [nA, nA + nB) and tile map big_tiles[nA + 1 .. nA + nB + 1] (each map
  // starts at 0). Nothing is written past big_cap segments or big_cap + 2
  // map entries; the host refuses more than big_cap oversized buckets.
  // (six scans in three barrier pairs; a total is the last thread's
  // exclusive value plus its own, broadcast through LDS)
  uint32_t ia, ib, ta, tbb, kx, nx;
  block_excl_scan2<1024>(nbig_a, nbig_b, tmp, &ia, &ib);
  block_excl_scan2<1024>(tiles_a, tiles_b, tmp, &ta, &tbb);
  block_excl_scan2<1024>(big_keys, nonempty, tmp, &kx, &nx);
  __shared__ uint32_t s_tot[6];
  if (tid == 1023) {
    s_tot[0] = ia + nbig_a;
    s_tot[1] = ib + nbig_b;
    s_tot[2] = ta + tiles_a;
    s_tot[3] = tbb + tiles_b;
    s_tot[4] = kx + big_keys;
    s_tot[5] = nx + nonempty;
  }
  __syncthreads();
  const uint32_t na_all = s_tot[0], nb_all = s_tot[1], ta_all = s_tot[2], tb_all = s_tot[3],
                 keys_all = s_tot[4], ne_all = s_tot[5];
  const uint32_t cap = static_cast<uint32_t>(tb.big_cap);
  const uint32_t map_b = na_all + 1u;  // first map entry of group B
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    if (v[q] > static_cast<uint32_t>(kLearnBigCap)) {
      const int32_t b = tid * PER + q;
      const uint32_t tiles = (v[q] + fallback_tile - 1) / fallback_tile;
      const bool in_b = (wide >> q) & 1u;
      const uint32_t seg = in_b ? na_all + ib : ia;
      const uint32_t map = in_b ? map_b + ib : ia;
      if (seg < cap && map < cap + 2u) {
        tb.big_start[seg] = s_bs[b];
        tb.big_end[seg] = s_bs[b + 1];
        tb.big_tiles[map] = static_cast<int32_t>(in_b ? tbb : ta);
      }
      if (in_b) {
        tbb += tiles;
        ++ib;
      } else {
        ta += tiles;
        ++ia;
      }
    }
  }
  const uint32_t n_big_all = na_all + nb_all;
  if (tid == 0) {
    if (na_all < cap + 2u) tb.big_tiles[na_all] = static_cast<int32_t>(ta_all);
    if (map_b + nb_all < cap + 2u) tb.big_tiles[map_b + nb_all] = static_cast<int32_t>(tb_all);
  }
  // coarse segments: counts, starts, tile map; fine counts per segment
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = tid * PER + q;
    tb.hist_b[(i / kLearnFine) * 1024 + (i % kLearnFine)] = v[q];
  }
  uint32_t cnt = 0;
  if (tid < kLearnCoarse) {
    cnt = static_cast<uint32_t>(s_bs[(tid + 1) * kLearnFine] - s_bs[tid * kLearnFine]);
    tb.hist_a[tid] = cnt;
    tb.seg_start[tid] = s_bs[tid * kLearnFine];
  }
  const uint32_t tiles = (cnt + kLearnPassTile - 1) / kLearnPassTile;
  uint32_t ttot;
  const uint32_t tex = block_excl_scan<1024>(tiles, tmp, &ttot);
  if (tid < kLearnCoarse) tb.tile_start[tid] = static_cast<int32_t>(tex);
  if (tid == 0) {
    tb.seg_start[kLearnCoarse] = static_cast<int32_t>(total);
    tb.tile_start[kLearnCoarse] = static_cast<int32_t>(ttot);
    *tb.tiles_b = static_cast<int32_t>(ttot);
    tb.info[kInfoBig] = static_cast<int32_t>(n_big_all);
    tb.info[kInfoBigTiles] = static_cast<int32_t>(ta_all + tb_all);
    tb.info[kInfoBigA] = static_cast<int32_t>(na_all);
    tb.info[kInfoTilesA] = static_cast<int32_t>(ta_all);
    tb.info[kInfoBigKeys] = static_cast<int32_t>(keys_all);
    tb.info[kInfoMaxBucket] = s_max;
    tb.info[kInfoNonEmpty] = static_cast<int32_t>(ne_all);
    tb.info[kInfoVary] = static_cast<int32_t>(s_vary);
  }
}

// ------------------------------------------------------------ bucket sort
// One workgroup per bucket whose size is in (lo_cap, WAVES * 64 * KPT].
template <int WAVES, int KPT>
__global__ __launch_bounds__(WAVES * 64) void lsort_kernel(uint32_t* __restrict__ keys,
                                                           const int32_t* __restrict__ bstart,
                                                           int32_t lo_cap) {
  const int32_t b = blockIdx.x;
  const int32_t lo = bstart[b];
  const int32_t n = bstart[b + 1] - lo;
  if (n <= lo_cap || n > WAVES * 64 * KPT) return;
  lds_bucket_sort<WAVES, KPT>(keys + lo, n);
}

// The last key of each non-empty bucket must not exceed the next key.
__global__ __launch_bounds__(256) void lverify_kernel(const uint32_t* __restrict__ keys,
                                                      const int32_t* __restrict__ bstart,
                                                      int32_t n, uint32_t* __restrict__ ctrl) {
  const int32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= kLearnBuckets) return;
  const int32_t s = bstart[b], e = bstart[b + 1];
  if (e > s && e < n && keys[e - 1] > keys[e]) atomicAdd(&ctrl[kCtrlLearnBad], 1u);
}

// ------------------------------------------------------------ training
__global__ __launch_bounds__(256) void lsample_kernel(const uint32_t* __restrict__ keys,
                                                      int32_t stride,
                                                      uint32_t* __restrict__ ring,
                                                      int64_t ring_cap, int64_t ring_pos,
                                                      int64_t count) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < count) ring[(ring_pos + i) % ring_cap] = keys[i * stride];
}

__global__ void lmeta_kernel(const uint32_t* __restrict__ sorted, int32_t ns,
                             uint32_t* __restrict__ meta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    meta[0] = sorted[0];
    meta[1] = sorted[ns - 1];
  }
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Mini-batch SGD with momentum on the mean of 0.5 * (r(x) - t)^2, sample i
// of the sorted samples having target t = (i + 0.5) / ns. One workgroup of
// 1024 threads, one sample per thread per step; the gradient is summed over
// the workgroup in a fixed order, so training is deterministic.
__global__ __launch_bounds__(1024) void ltrain_root_kernel(const uint32_t* __restrict__ sorted,
                                                           int32_t ns,
                                                           const uint32_t* __restrict__ meta,
                                                           int32_t steps, float lr,
                                                           float momentum, uint32_t seed,
                                                           float* __restrict__ model,
                                                           float* __restrict__ loss_out) {
  constexpr int P = kCdfRootParams;
  __shared__ float s_m[P];
  __shared__ float s_v[P];
  __shared__ float s_part[16][P + 1];
  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  const CdfDev cv = cdf_view(model, meta, nullptr);
  if (tid < P) {
    s_m[tid] = model[tid];
    s_v[tid] = 0.0f;
  }
  __syncthreads();
  const float inv_ns = 1.0f / static_cast<float>(ns);
  float last_loss = 0.0f;
  for (int32_t step = 0; step < steps; ++step) {
    const uint32_t idx =
        hash32(seed ^ hash32(static_cast<uint32_t>(step) * 1024u + static_cast<uint32_t>(tid))) %
        static_cast<uint32_t>(ns);
    const float x = cdf_x(cv, sorted[idx]);
    const float t = (static_cast<float>(idx) + 0.5f) * inv_ns;
    float h[kCdfHidden];
    float r = s_m[kCdfB2];
#pragma unroll
    for (int j = 0; j < kCdfHidden; ++j) {
      h[j] = fmaxf(s_m[kCdfW1 + j] * x + s_m[kCdfB1 + j], 0.0f);
      r = r + s_m[kCdfW2 + j] * h[j];
    }
    const float err = r - t;
    last_loss = 0.5f * err * err;
    float gr[P];
#pragma unroll
    for (int j = 0; j < kCdfHidden; ++j) {
      const float dh = h[j] > 0.0f ? err * s_m[kCdfW2 + j] : 0.0f;
      gr[kCdfW1 + j] = dh * x;
      gr[kCdfB1 + j] = dh;
      gr[kCdfW2 + j] = err * h[j];
    }
    gr[kCdfB2] = err;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float g = gr[p];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
      if (lane == 0) s_part[w][p] = g;
    }
    __syncthreads();
    if (tid < P) {
      float g = 0.0f;
#pragma unroll
      for (int q = 0; q < 16; ++q) g += s_part[q][tid];
      g *= (1.0f / 1024.0f);
      s_v[tid] = momentum * s_v[tid] + g;
      s_m[tid] = s_m[tid] - lr * s_v[tid];
    }
    __syncthreads();
    if (tid < kCdfHidden && s_m[kCdfW1 + tid] * s_m[kCdfW2 + tid] < 0.0f) {
      s_m[kCdfW2 + tid] = 0.0f;
      s_v[kCdfW2 + tid] = 0.0f;
    }
    __syncthreads();
  }
  if (tid < P) model[tid] = s_m[tid];
  float l = last_loss;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if (lane == 0) s_part[w][0] = l;
  __syncthreads();
  if (tid == 0) {
    float s = 0.0f;
    for (int q = 0; q < 16; ++q) s += s_part[q][0];
    *loss_out = s * (1.0f / 1024.0f);
  }
}

// K_l = min{key : l(key) >= l} by binary search over the key space
// (l(key) is non-decreasing and l(0xffffffff) = 4095); K_4096 = 0xffffffff.
__global__ __launch_bounds__(256) void lleaf_starts_kernel(const float* __restrict__ model,
                                                           const uint32_t* __restrict__ meta,
                                                           uint32_t* __restrict__ kleaf) {
  const int32_t l = blockIdx.x * 256 + threadIdx.x;
  if (l > kCdfLeaves) return;
  if (l == 0 || l == kCdfLeaves) {
    kleaf[l] = l == 0 ? 0u : 0xffffffffu;
    return;
  }
  const CdfDev v = cdf_view(model, meta, nullptr);
  uint32_t lo = 0, hi = 0xffffffffu;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (leaf_of(v, mid) >= l) hi = mid;
    else lo = mid + 1;
  }
  kleaf[l] = lo;
}

__device__ __forceinline__ int32_t lower_bound_u32(const uint32_t* a, int32_t n, uint32_t key) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Leaf l: lo_l = the sample CDF at K_l; (a_l, b_l) = least squares of
// t = a + b * (key - K_l) over the samples in [K_l, K_{l+1}), b >= 0.
__global__ __launch_bounds__(256) void lfit_leaves_kernel(const uint32_t* __restrict__ sorted,
                                                          int32_t ns,
                                                          const uint32_t* __restrict__ kleaf,
                                                          float* __restrict__ model) {
  const int32_t l = blockIdx.x * 256 + threadIdx.x;
  if (l == 0) model[kCdfLeafLo + kCdfLeaves] = 1.0f;
  if (l >= kCdfLeaves) return;
  const uint32_t k0 = kleaf[l];
  const int32_t s = l == 0 ? 0 : lower_bound_u32(sorted, ns, k0);
  const int32_t e = l == kCdfLeaves - 1 ? ns : lower_bound_u32(sorted, ns, kleaf[l + 1]);
  const double inv = 1.0 / ns;
  double sd = 0, st = 0, sdd = 0, sdt = 0;
  for (int32_t i = s; i < e; ++i) {
    const double d = static_cast<double>(sorted[i] - k0);
    const double t = (i + 0.5) * inv;
    sd += d;
    st += t;
    sdd += d * d;
    sdt += d * t;
  }
  const double cnt = e - s;
  double a = s * inv, b = 0.0;
  if (cnt >= 1) a = st / cnt;
  if (cnt >= 2) {
    const double den = cnt * sdd - sd * sd;
    if (den > 0.0) {
      b = (cnt * sdt - sd * st) / den;
      if (b < 0.0) b = 0.0;
      a = (st - b * sd) / cnt;
    }
  }
  model[kCdfLeafLo + l] = static_cast<float>(s * inv);
  model[kCdfLeafA + l] = static_cast<float>(a);
  model[kCdfLeafB + l] = static_cast<float>(b);
}

// Linear-CDF leaves for the current root: y runs linearly from l / 4096 at
// K_l to (l + 1) / 4096 at K_{l+1} (the last leaf ends at kmax + 1).
__global__ __launch_bounds__(256) void llinear_leaves_kernel(const uint32_t* __restrict__ kleaf,
                                                             const uint32_t* __restrict__ meta,
                                                             float* __restrict__ model) {
  const int32_t l = blockIdx.x * 256 + threadIdx.x;
  if (l > kCdfLeaves) return;
  model[kCdfLeafLo + l] = static_cast<float>(l) / static_cast<float>(kCdfLeaves);
  if (l == kCdfLeaves) return;
  const uint64_t k0 = kleaf[l];
  const uint64_t k1 = l == kCdfLeaves - 1 ? static_cast<uint64_t>(meta[1]) + 1u : kleaf[l + 1];
  const double span = k1 > k0 ? static_cast<double>(k1 - k0) : 1.0;
  model[kCdfLeafA + l] = static_cast<float>(l) / static_cast<float>(kCdfLeaves);
  model[kCdfLeafB + l] = static_cast<float>(1.0 / (kCdfLeaves * span));
}

// Identity root r(x) = x, the start of training: units 0..7 are
// relu(x - j/8) (convex kinks), units 8..15 relu((j - 7.5)/8 - x) (concave
// kinks once their output weight turns negative); only unit 0 has a
// non-zero output weight. Units of both directions are needed: with
// w1 * w2 >= 0, units with w1 > 0 can only add convex kinks.
__global__ void llinear_root_kernel(float* __restrict__ model) {
  const int j = threadIdx.x;
  if (j < kCdfHidden) {
    const bool up = j < kCdfHidden / 2;
    model[kCdfW1 + j] = up ? 1.0f : -1.0f;
    model[kCdfB1 + j] = up ? -static_cast<float>(j) / 8.0f : (static_cast<float>(j) - 7.5f) / 8.0f;
    model[kCdfW2 + j] = j == 0 ? 1.0f : 0.0f;
  }
  if (j == 0) model[kCdfB2] = 0.0f;
}

// fb[f] = min{key : q(key) >= f}, then a running maximum so that the
// boundaries are non-decreasing whatever the parameters are.
__global__ __launch_bounds__(256) void lboundaries_kernel(const float* __restrict__ model,
                                                          const uint32_t* __restrict__ meta,
                                                          const uint32_t* __restrict__ kleaf,
                                                          uint32_t* __restrict__ fb) {
  const int32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= kLearnBuckets) return;
  if (f == 0) {
    fb[0] = 0u;
    return;
  }
  const CdfDev v = cdf_view(model, meta, kleaf);
  uint32_t lo = 0, hi = 0xffffffffu;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (q_model(v, mid) >= static_cast<uint32_t>(f)) hi = mid;
    else lo = mid + 1;
  }
  fb[f] = lo;
}

__global__ __launch_bounds__(1024) void lmonotone_kernel(uint32_t* __restrict__ fb) {
  __shared__ uint32_t s_m[1024];
  constexpr int PER = kLearnBuckets / 1024;
  const int tid = threadIdx.x;
  uint32_t v[PER];
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    v[q] = fb[tid * PER + q];
    m = max(m, v[q]);
  }
  s_m[tid] = m;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint32_t x = s_m[i];
      s_m[i] = run;
      run = max(run, x);
    }
  }
  __syncthreads();
  uint32_t run = s_m[tid];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    run = max(run, v[q]);
    fb[tid * PER + q] = run;
  }
}

// Cell table of the compiled boundaries: cshift so that 8192 cells of
// 2^cshift keys cover [kmin, kmax], then the bucket of every cell's first
// key (2^32 - 1 past the end of the key space).
__global__ __launch_bounds__(256) void lcells_kernel(const uint32_t* __restrict__ fb,
                                                     uint32_t* __restrict__ meta,
                                                     uint16_t* __restrict__ ctab) {
  const int32_t c = blockIdx.x * 256 + threadIdx.x;
  const uint32_t kmin = meta[0], span = meta[1] - meta[0];
  const uint32_t bits = span ? 32u - static_cast<uint32_t>(__clz(static_cast<int>(span))) : 0u;
  const uint32_t cshift = bits > static_cast<uint32_t>(kLearnCellBits)
                              ? bits - static_cast<uint32_t>(kLearnCellBits) : 0u;
  if (c == 0) meta[2] = cshift;
  if (c > kLearnCells) return;
  const uint64_t k64 = static_cast<uint64_t>(kmin) + (static_cast<uint64_t>(c) << cshift);
  const uint32_t key = k64 > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(k64);
  uint32_t f = 0;
  for (uint32_t step = kLearnBuckets / 2; step > 0; step >>= 1) f += fb[f + step] <= key ? step : 0u;
  ctab[c] = static_cast<uint16_t>(f);
}

// ------------------------------------------------------------ diagnostics
// Mean gradient of 0.5 * (r(x) - t)^2 over n pairs, accumulated into grad
// (zeroed by the caller).
__global__ __launch_bounds__(256) void lgrad_kernel(const float* __restrict__ xs,
                                                    const float* __restrict__ ts, int32_t n,
                                                    const float* __restrict__ model,
                                                    float* __restrict__ grad) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float x = xs[i], t = ts[i];
  float h[kCdfHidden];
  float r = model[kCdfB2];
#pragma unroll
  for (int j = 0; j < kCdfHidden; ++j) {
    h[j] = fmaxf(model[kCdfW1 + j] * x + model[kCdfB1 + j], 0.0f);
    r = r + model[kCdfW2 + j] * h[j];
  }
  const float err = r - t;
  const float sc = 1.0f / static_cast<float>(n);
#pragma unroll
  for (int j = 0; j < kCdfHidden; ++j) {
    const float dh = h[j] > 0.0f ? err * model[kCdfW2 + j] : 0.0f;
    atomicAdd(&grad[kCdfW1 + j], dh * x * sc);
    atomicAdd(&grad[kCdfB1 + j], dh * sc);
    atomicAdd(&grad[kCdfW2 + j], err * h[j] * sc);
  }
  atomicAdd(&grad[kCdfB2], err * sc);
}

__global__ __launch_bounds__(256) void lbucket_kernel(const uint32_t* __restrict__ keys,
                                                      int32_t n, const float* __restrict__ model,
                                                      const uint32_t* __restrict__ meta,
                                                      const uint32_t* __restrict__ kleaf,
                                                      const uint32_t* __restrict__ fb,
                                                      const uint16_t* __restrict__ ctab, int mode,
                                                      uint32_t* __restrict__ q) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = keys[i];
  if (mode == 0) {
    q[i] = q_model(cdf_view(model, meta, kleaf), key);
  } else if (mode == 2) {
    uint32_t lo, hi;
    cell_bounds(ctab, meta[0], meta[2], key, lo, hi);
    q[i] = search_range(fb, 0u, lo, hi, key);
  } else {
    uint32_t f = 0;
    for (uint32_t step = kLearnBuckets / 2; step > 0; step >>= 1)
      f += fb[f + step] <= key ? step : 0u;
    q[i] = f;
  }
}

}  // namespace

// ------------------------------------------------------------ launchers
void launch_lhist(const uint32_t* keys, int64_t n, const LearnTables& tb, int32_t* n_rows,
                  hipStream_t s) {
  const int64_t want = (n + 65535) / 65536;
  const int32_t g = static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(want, kLearnHistGrid)));
  *n_rows = g;
  lhist_kernel<<<g, kHistThreads, 0, s>>>(keys, n, tb, tb.rows);
  lreduce_kernel<<<dim3(kLearnBuckets / 256, kLearnReduceGroups), 256, 0, s>>>(tb.rows, g, tb.hist);
  lreduce2_kernel<<<kLearnBuckets / 256, 256, 0, s>>>(tb.hist);
}

void launch_lprefix(const LearnTables& tb, int32_t, int32_t fallback_tile, hipStream_t s) {
  lprefix_kernel<<<1, 1024, 0, s>>>(tb, fallback_tile);
}

void launch_lpass_a(const uint32_t* in, uint32_t* out, int32_t n, const LearnTables& tb,
                    uint64_t*, uint32_t* ctrl, hipStream_t s) {
  if (n <= 0) return;
  SegView sv{};
  sv.n_seg = 1;
  sv.n1 = n;
  const int32_t tiles = (n + kLearnPassTile - 1) / kLearnPassTile;
  cursor_scatter_kernel<kLearnPassWaves * 64, kLearnPassKpt, CoarseDigit>
      <<<tiles, kLearnPassWaves * 64, 0, s>>>(in, out, sv, nullptr, nullptr, tb.hist_a, 0,
                                               tb.cursor, ctrl, CoarseDigit{tb});
}

int32_t learn_pass_b_max_tiles(int32_t n) {
  return (n + kLearnPassTile - 1) / kLearnPassTile + kLearnCoarse;
}

void launch_lpass_b(const uint32_t* in, uint32_t* out, int32_t n, const LearnTables& tb,
                    uint64_t*, uint32_t* ctrl, hipStream_t s) {
  if (n <= 0) return;
  SegView sv{};
  sv.n_seg = kLearnCoarse;
  sv.start = tb.seg_start;
  sv.n1 = n;
  cursor_scatter_kernel<kLearnPassWaves * 64, kLearnPassKpt, FineDigit>
      <<<learn_pass_b_max_tiles(n), kLearnPassWaves * 64, 0, s>>>(
          in, out, sv, tb.tile_start, tb.tiles_b, tb.hist_b, 1024, tb.cursor + kLearnCoarse, ctrl,
          FineDigit{tb});
}

void launch_lsort(uint32_t* keys, const LearnTables& tb, hipStream_t s) {
  lsort_kernel<4, 16><<<kLearnBuckets, 256, 0, s>>>(keys, tb.bstart, 1);
  lsort_kernel<8, 16><<<kLearnBuckets, 512, 0, s>>>(keys, tb.bstart, kLearnSmallCap);
  // 512 x 24 keys: two workgroups per CU by LDS; 16384 keys take one per CU
  lsort_kernel<8, 24><<<kLearnBuckets, 512, 0, s>>>(keys, tb.bstart, kLearnMidCap);
  lsort_kernel<16, 16><<<kLearnBuckets, 1024, 0, s>>>(keys, tb.bstart, kLearnLargeCap);
  static_assert(8 * 64 * 24 == kLearnLargeCap, "bucket-sort classes");
  static_assert(16 * 64 * 16 == kLearnBigCap, "largest bucket-sort class");
}

void launch_lverify(const uint32_t* keys, const LearnTables& tb, int32_t n, uint32_t* ctrl,
                    hipStream_t s) {
  lverify_kernel<<<kLearnBuckets / 256, 256, 0, s>>>(keys, tb.bstart, n, ctrl);
}

void launch_lsample(const uint32_t* keys, int64_t, int32_t stride, uint32_t* ring,
                    int64_t ring_cap, int64_t ring_pos, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  lsample_kernel<<<static_cast<int>((count + 255) / 256), 256, 0, s>>>(keys, stride, ring,
                                                                      ring_cap, ring_pos, count);
}

void launch_lmeta(const uint32_t* sorted, int32_t ns, uint32_t* meta, hipStream_t s) {
  lmeta_kernel<<<1, 64, 0, s>>>(sorted, ns, meta);
}

void launch_ltrain_root(const uint32_t* sorted, int32_t ns, const uint32_t* meta, int32_t steps,
                        float lr, float momentum, uint32_t seed, float* model, float* loss,
                        hipStream_t s) {
  ltrain_root_kernel<<<1, 1024, 0, s>>>(sorted, ns, meta, steps, lr, momentum, seed, model, loss);
}

void launch_lleaf_starts(const float* model, const uint32_t* meta, uint32_t* kleaf,
                         hipStream_t s) {
  lleaf_starts_kernel<<<(kCdfLeaves + 1 + 255) / 256, 256, 0, s>>>(model, meta, kleaf);
}

void launch_lfit_leaves(const uint32_t* sorted, int32_t ns, const uint32_t* kleaf, float* model,
                        hipStream_t s) {
  lfit_leaves_kernel<<<kCdfLeaves / 256, 256, 0, s>>>(sorted, ns, kleaf, model);
}

void launch_lboundaries(const float* model, uint32_t* meta, const uint32_t* kleaf,
                        uint32_t* fb, uint16_t* ctab, hipStream_t s) {
  lboundaries_kernel<<<kLearnBuckets / 256, 256, 0, s>>>(model, meta, kleaf, fb);
  lmonotone_kernel<<<1, 1024, 0, s>>>(fb);
  lcells_kernel<<<(kLearnCells + 1 + 255) / 256, 256, 0, s>>>(fb, meta, ctab);
}

void launch_llinear(float* model, hipStream_t s) { llinear_root_kernel<<<1, 64, 0, s>>>(model); }

void launch_llinear_leaves(const uint32_t* kleaf, const uint32_t* meta, float* model,
                           hipStream_t s) {
  llinear_leaves_kernel<<<(kCdfLeaves + 1 + 255) / 256, 256, 0, s>>>(kleaf, meta, model);
}

void launch_lgrad(const float* xs, const float* ts, int32_t n, const float* model, float* grad,
                  hipStream_t s) {
  if (n > 0) lgrad_kernel<<<(n + 255) / 256, 256, 0, s>>>(xs, ts, n, model, grad);
}

void launch_lbucket(const uint32_t* keys, int32_t n, const float* model, const uint32_t* meta,
                    const uint32_t* kleaf, const uint32_t* fb, const uint16_t* ctab, int mode,
                    uint32_t* q, hipStream_t s) {
  if (n > 0)
    lbucket_kernel<<<(n + 255) / 256, 256, 0, s>>>(keys, n, model, meta, kleaf, fb, ctab, mode, q);
}

}  // namespace ocp

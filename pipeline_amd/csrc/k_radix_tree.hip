// Stage 4: Karras binary radix tree over the sorted unique codes, one
// thread per node, exactly the per-node function of
// kernels/04_radix_tree.cpp:47-120 (kernels/cuda/04_radix_tree.cu:33-113):
// direction from the two neighbours, exponential then binary search for
// the far end j, common-prefix length delta(i, j) = clz(a ^ b) - 1, binary
// search for the split gamma, leaf flags, parent links of internal
// children. clz runs on v_ffbh_u32 with clz(0) := 32.
//
// Parents without a race. The reference writes parent[] from every node;
// run sequentially the last writer wins. Nodes i >= 1 compute the standard
// Karras ranges over all U keys, so every internal node except the true
// root's children has exactly one writer among them. Node 0 is the
// quirk root: the reference gives it the range [0, n-1] instead of [0, n]
// (04_radix_tree.cpp:63-65), and it writes first, so whatever it writes is
// either overwritten by a node i >= 1 or lands on an entry nobody else
// writes, which the true root would own. Thread 0 therefore computes the
// true root split and stores 0 into parent[0] and into the parent entries
// of the true root's internal children; nodes i >= 1 store plainly. The
// result equals the sequential reference on zeroed memory without atomics
// or a zero-fill pass. With the same writes the fused build also stores
// prefix_n[parent[i]] at parent_pn[i], so the edge count needs no gather.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

__device__ __forceinline__ uint32_t clz32(uint32_t x) {
  return x ? static_cast<uint32_t>(__clz(static_cast<int>(x))) : 32u;
}
__device__ __forceinline__ uint32_t delta(uint32_t a, uint32_t b) {
  return (clz32(a ^ b) - 1u) & 0xffu;
}
__device__ __forceinline__ int32_t log2_ceil(uint32_t x) {
  if (x == 0) return 0;
  const int32_t lo = 31 - static_cast<int32_t>(clz32(x));
  return lo + ((1u << lo) < x);
}

// Codes of a segment, served from an LDS window [lo, lo + n) when the
// index falls inside it. Most searches of a node stay close to the node.
struct Codes {
  const uint32_t* g;
  lds_ptr<const uint32_t> w;
  int32_t lo;
  int32_t n;
  __device__ __forceinline__ uint32_t operator[](int32_t q) const {
    return static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n) ? w[q - lo] : g[q];
  }
  // For an index known to lie in the window: no fallback branch.
  __device__ __forceinline__ uint32_t local(int32_t q) const { return w[q - lo]; }
};

// delta(code_i, c) > dl as one subtraction and compare, for a fixed dl in
// [0, 30]. With x = code_i ^ c: x == 0 gives delta 31 and x >= 2^31 gives
// delta 255 (clz 0, minus one, as a byte), both above dl; otherwise delta =
// clz(x) - 1 > dl exactly when x < 2^(30 - dl). So the predicate holds for
// x < lim or x >= 2^31, i.e. for x - lim >= 2^31 - lim (unsigned).
struct Above {
  uint32_t ci;
  uint32_t lim;
  uint32_t thr;
  __device__ __forceinline__ Above(uint32_t code_i, uint32_t dl)
      : ci(code_i), lim(1u << (30u - dl)), thr(0x80000000u - (1u << (30u - dl))) {}
  __device__ __forceinline__ bool operator()(uint32_t c) const { return ((c ^ ci) - lim) >= thr; }
};

#ifndef OCP_RADIX_HALO
#define OCP_RADIX_HALO 1024
#endif
// Codes staged on either side of a tile. 10M uniform keys: 128 -> 175 us,
// 256 -> 141 us, 512 -> 131 us (the neighbours' halos mostly hit in L2);
// with the wave path for wide nodes, 512 -> 110 us, 1024 -> 105 us,
// 2048 -> 109 us (2^24 clustered keys: 171, 160, 169 us).
constexpr int kHalo = OCP_RADIX_HALO;
// Nodes whose exponential search is still running at this length are
// finished by a whole wave; the rest stay inside the LDS window.
constexpr int32_t kLocalMax = kHalo / 2;
constexpr int kLocalLog = __builtin_ctz(kLocalMax);
static_assert((kLocalMax & (kLocalMax - 1)) == 0 && kLocalLog >= 1 && kLocalLog < 31,
              "the local search bound is a power of two");

// Split search of 04_radix_tree.cpp:94-104 for node i with far end j. The
// reference's divisor runs over 2, 4, ..., 2^log2_ceil(l) and its step is
// ceil(l / divisor); with power-of-two divisors that is a shift, so no
// integer division is left on the device.
// Only called for local nodes (l < kLocalMax), whose probes stay in the window.
template <bool CHECK>
__device__ __forceinline__ int32_t split_search(const Codes& codes, const Above& above,
                                                int32_t i, int32_t d,
                                                int32_t l, int32_t n) {
  int32_t s = 0;
  const int32_t lg = log2_ceil(static_cast<uint32_t>(l));
  const int32_t s_cutoff = (d == -1) ? i - 1 : n - i - 1;
  // l < kLocalMax, so lg <= kLocalLog: a fixed number of guarded steps.
#pragma unroll
  for (int32_t k = 1; k <= kLocalLog; ++k) {
    const int32_t t = static_cast<int32_t>((static_cast<uint32_t>(l) + (1u << k) - 1u) >> k);
    const int32_t st = s + t;
    const bool probe = k <= lg && (!CHECK || st <= s_cutoff);
    if (probe && above(codes.local(probe ? i + st * d : i))) s = st;
  }
  return i + s * d + min(d, 0);
}

// Wide nodes. All three searches of a node look for the largest x with a
// predicate that is monotone in x (codes are sorted, so the common prefix
// with code i only shrinks away from i): the reference's exponential search
// stops at the first power of two that fails, its binary searches return
// the largest x below their bound that holds (for the length: x < l_max and
// x <= l_cutoff; for the split: x <= min(l - 1, s_cutoff), the predicate
// failing at l). A whole wave evaluates 64 candidates per round trip
// instead of one, so a node spanning 2^k keys costs about k/6 dependent
// loads instead of 3k.
template <class Pred>
__device__ __forceinline__ int32_t wave_max_true(int32_t lo, int32_t hi, Pred pred) {
  // largest x in [lo, hi) with pred(x); pred(lo) holds.
  const int32_t lane = static_cast<int32_t>(lane_id());
  while (hi - lo > 1) {
    const int32_t step = (hi - lo + 63) >> 6;
    const int32_t x = lo + step * (lane + 1);
    const bool ok = x < hi && pred(x);
    const int32_t c = __popcll(__ballot(ok));
    const int32_t nlo = lo + step * c;
    hi = min(hi, lo + step * (c + 1));
    lo = nlo;
  }
  return lo;
}

struct NodeOut {
  uint32_t delta_node;
  int32_t gamma;
  int32_t j;
};

__device__ __forceinline__ NodeOut wave_node(const Codes& codes, int32_t i, int32_t n) {
  const int32_t lane = static_cast<int32_t>(lane_id());
  const uint32_t code_i = codes[i];
  int32_t d, l;
  if (i == 0) {
    d = 1;
    l = n - 1;
  } else {
    const int32_t diff = static_cast<int32_t>(delta(code_i, codes[i + 1])) -
                         static_cast<int32_t>(delta(code_i, codes[i - 1]));
    d = (diff > 0) - (diff < 0);
    const uint32_t delta_min = delta(code_i, codes[i - d]);
    // exponential search: lane k tests l_max = 2^k, k = 1..30
    bool ok = false;
    if (lane >= 1 && lane <= 30) {
      const int64_t q = static_cast<int64_t>(i) + (int64_t(1) << lane) * d;
      ok = q >= 0 && q <= n && delta(code_i, codes[static_cast<int32_t>(q)]) > delta_min;
    }
    const int32_t c = __popcll(__ballot(ok));
    const int32_t l_max = 1 << (c + 1);
    const int32_t l_cutoff = (d == -1) ? i : n - i;
    l = wave_max_true(0, min(l_max, l_cutoff + 1), [&](int32_t x) {
      return delta(code_i, codes[i + x * d]) > delta_min;
    });
  }
  NodeOut o;
  o.j = i + l * d;
  o.delta_node = delta(code_i, codes[o.j]);
  const int32_t s_cutoff = (d == -1) ? i - 1 : n - i - 1;
  const int32_t s = l <= 0 ? 0 : wave_max_true(0, min(l - 1, s_cutoff) + 1, [&](int32_t x) {
    return delta(code_i, codes[i + x * d]) > o.delta_node;
  });
  o.gamma = i + s * d + min(d, 0);
  return o;
}


// A node whose searches stay within kLocalMax of it, read from the LDS
// window without the window test (the window spans kHalo >= kLocalMax + 2
// around the tile). Returns false, emitting nothing, for a node the wave
// path must finish: a far end beyond kLocalMax, or a threshold of 31 or
// above (repeated codes, codes with bit 31 set), for which the wave path
// evaluates delta itself. CHECK: bound the probes by [0, n] and the
// reference's length and split cutoffs (tiles near either end).
template <bool CHECK, class Emit>
__device__ __forceinline__ bool local_node(const Codes& codes, int32_t i, int32_t n,
                                           Emit& emit) {
  const uint32_t code_i = codes.local(i);
  const uint32_t dn = delta(code_i, codes.local(i + 1));
  const uint32_t dp = delta(code_i, codes.local(i - 1));
  const int32_t diff = static_cast<int32_t>(dn) - static_cast<int32_t>(dp);
  const int32_t d = (diff > 0) - (diff < 0);
  // delta(code_i, codes[i - d]); d == 0 (equal neighbours, repeated codes
  // only) compares code i with itself: 31.
  const uint32_t delta_min = d > 0 ? dp : (d < 0 ? dn : 31u);
  if (delta_min >= 31u) return false;
  const Above above_min(code_i, delta_min);
  // Exponential search for an upper bound l_max (a power of two), then
  // binary search of the length: t = l_max / 2, l_max / 4, ..., 1. The
  // reference probes 2, 4, 8, ... until one fails; here the probes up to
  // kLocalMax are read at once and l_max = 2^(leading passes + 1).
  uint32_t pass = 0;
#pragma unroll
  for (int k = 1; k <= kLocalLog; ++k) {
    const int32_t q = i + (1 << k) * d;
    const bool in = !CHECK || (q >= 0 && q <= n);
    const uint32_t cq = codes.local(in ? q : i);
    pass |= static_cast<uint32_t>(in && above_min(cq)) << (k - 1);
  }
  const int32_t l_max = 2 << __builtin_ctz(~pass);
  if (l_max > kLocalMax) return false;
  const int32_t l_cutoff = (d == -1) ? i : n - i;
  int32_t l = 0;
#pragma unroll
  for (int k = 1; k <= kLocalLog; ++k) {  // t = l_max / 2, l_max / 4, ..., 1
    const int32_t t = l_max >> k;
    const bool probe = t >= 1 && (!CHECK || l + t <= l_cutoff);
    if (probe && above_min(codes.local(probe ? i + (l + t) * d : i))) l += t;
  }
  const int32_t j = i + l * d;
  const uint32_t delta_node = delta(code_i, codes.local(j));
  if (delta_node >= 31u) return false;
  emit(i, delta_node, split_search<CHECK>(codes, Above(code_i, delta_node), i, d, l, n), j);
  return true;
}

__global__ __launch_bounds__(256) void radix_tree_kernel(
    const uint32_t* __restrict__ codes_all, uint8_t* __restrict__ prefix_all,
    uint8_t* __restrict__ hl_all, uint8_t* __restrict__ hr_all,
    int32_t* __restrict__ lc_all, int32_t* __restrict__ par_all,
    uint8_t* __restrict__ ppn_all, SegView sv,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ n_unique,
    int32_t n_brt_host, const uint32_t* __restrict__ ctrl) {
  static_assert(kLocalMax <= kHalo / 2 + 1, "local searches must stay in the window");
  __shared__ uint32_t s_codes[kNodeTile + 2 * kHalo];
  __shared__ int32_t s_wide[kNodeTile];
  __shared__ uint32_t s_nwide;
  const TileRef r = find_tile(tile_start, sv.n_seg, xcd_tile(blockIdx.x, gridDim.x));
  const int32_t n = n_unique ? n_unique[r.seg] - 1 : n_brt_host;
  const int32_t sb = seg_begin(sv, r.seg);  // loaded with n, before the test
  const int32_t i0 = r.local * kNodeTile;
  // After a lookback time-out the unique keys are not valid input.
  if (i0 >= n || (ctrl && stopped_any(ctrl))) return;
  int32_t* parent = par_all + sb;
  uint8_t* parent_pn = ppn_all ? ppn_all + sb : nullptr;
  if (threadIdx.x == 0) s_nwide = 0;
  // Stage codes [i0 - halo, i0 + kNodeTile + halo) of the n+1 codes.
  const int32_t lo = max(0, i0 - kHalo);
  const int32_t hi = min(n + 1, i0 + kNodeTile + kHalo);
  {
    // every load of the window in flight at once
    constexpr int kRows = (kNodeTile + 2 * kHalo + 255) / 256;
    uint32_t rc[kRows];
    const uint32_t* src = codes_all + sb + lo + static_cast<int32_t>(threadIdx.x);
    if (hi - lo == kRows * 256) {  // a full window: no per-load guard
#pragma unroll
      for (int k = 0; k < kRows; ++k) rc[k] = src[k * 256];
    } else {
#pragma unroll
      for (int k = 0; k < kRows; ++k)
        if (lo + static_cast<int32_t>(threadIdx.x) + k * 256 < hi) rc[k] = src[k * 256];
    }
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
      const int32_t q = lo + static_cast<int32_t>(threadIdx.x) + k * 256;
      if (q < hi) s_codes[q - lo] = rc[k];
    }
  }
  __syncthreads();
  const Codes codes{codes_all + sb, as_lds<const uint32_t>(s_codes), lo, hi - lo};

  auto emit = [&](int32_t node, uint32_t delta_node, int32_t gamma, int32_t j) {
    const bool leaf_l = min(node, j) == gamma;
    const bool leaf_r = max(node, j) == gamma + 1;
    prefix_all[sb + node] = static_cast<uint8_t>(delta_node);
    hl_all[sb + node] = leaf_l;
    hr_all[sb + node] = leaf_r;
    lc_all[sb + node] = gamma;
    if (node > 0) {
      if (!leaf_l) parent[gamma] = node;
      if (!leaf_r) parent[gamma + 1] = node;
      if (parent_pn) {
        if (!leaf_l) parent_pn[gamma] = static_cast<uint8_t>(delta_node);
        if (!leaf_r) parent_pn[gamma + 1] = static_cast<uint8_t>(delta_node);
      }
    }
  };

  // A tile at least 2 * kLocalMax away from both ends of [0, n]: every
  // probe of a local node (within kLocalMax + 2 of i) is inside [0, n] and
  // below the length and split cutoffs, so those checks are skipped.
  const bool interior = i0 >= 2 * kLocalMax && i0 + kNodeTile + 2 * kLocalMax <= n;
  for (int32_t k = 0; k < kNodeTile / 256; ++k) {
  const int32_t i = i0 + k * 256 + static_cast<int32_t>(threadIdx.x);
  bool wide = false;
  if (i < n) {
    if (i == 0) {
      wide = true;
    } else {
      wide = interior ? !local_node<false>(codes, i, n, emit)
                      : !local_node<true>(codes, i, n, emit);
    }
  }
  const uint64_t bal = __ballot(wide);
  if (bal) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(&s_nwide, static_cast<uint32_t>(__popcll(bal)));
    base = __shfl(base, 0, 64);
    if (wide) s_wide[base + __popcll(bal & lanemask_lt())] = i;
  }
  }
  __syncthreads();
  const uint32_t nw = s_nwide;
  for (uint32_t q = wave_id(); q < nw; q += 4) {
    const int32_t node = s_wide[q];
    const NodeOut o = wave_node(codes, node, n);
    if (lane_id() == 0) emit(node, o.delta_node, o.gamma, o.j);
    if (node == 0) {
      // The true root spans [0, n]: its internal children have no other
      // writer, and sequentially they end up holding node 0. Its split is
      // the largest s <= n - 1 with delta(c0, codes[s]) > delta(c0, codes[n]).
      const uint32_t c0 = codes[0];
      const uint32_t dn = delta(c0, codes[n]);
      const int32_t g0 = wave_max_true(0, n, [&](int32_t q) { return delta(c0, codes[q]) > dn; });
      if (lane_id() == 0) {
        parent[0] = 0;
        if (g0 != 0 && g0 < n) parent[g0] = 0;
        if (g0 + 1 != n && g0 + 1 < n) parent[g0 + 1] = 0;
        if (parent_pn) {
          const uint8_t pn0 = static_cast<uint8_t>(o.delta_node);
          parent_pn[0] = pn0;
          if (g0 != 0 && g0 < n) parent_pn[g0] = pn0;
          if (g0 + 1 != n && g0 + 1 < n) parent_pn[g0 + 1] = pn0;
        }
      }
    }
  }
}

}  // namespace

void launch_radix_tree(const uint32_t* codes, uint8_t* prefix_n,
                       uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                       int32_t* left_child, int32_t* parent, uint8_t* parent_pn,
                       const SegView& sv, const int32_t* tile_start, int32_t total_tiles,
                       const int32_t* n_unique, int32_t n_brt_host,
                       const uint32_t* ctrl, hipStream_t s) {
  if (total_tiles <= 0) return;
  static_assert(kNodeTile % 256 == 0, "whole node rows per thread");
  radix_tree_kernel<<<total_tiles, 256, 0, s>>>(
      codes, prefix_n, has_leaf_left, has_leaf_right, left_child, parent, parent_pn, sv,
      tile_start, n_unique, n_brt_host, ctrl);
}

}  // namespace ocp

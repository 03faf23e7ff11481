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
// or a zero-fill pass.
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
__device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) {
  return (a + b - 1) / b;
}

// Codes of a segment, served from an LDS window [lo, lo + n) when the
// index falls inside it. Most searches of a node stay close to the node.
struct Codes {
  const uint32_t* g;
  const uint32_t* w;
  int32_t lo;
  int32_t n;
  __device__ __forceinline__ uint32_t operator[](int32_t q) const {
    return static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n) ? w[q - lo] : g[q];
  }
};

constexpr int kHalo = 256;

// Split search of 04_radix_tree.cpp:94-104 for node i with far end j.
__device__ __forceinline__ int32_t split_search(const Codes& codes,
                                                int32_t i, int32_t d,
                                                int32_t l, int32_t n,
                                                uint32_t delta_node) {
  const uint32_t code_i = codes[i];
  int32_t s = 0;
  const int64_t max_divisor = int64_t(1) << log2_ceil(static_cast<uint32_t>(l));
  const int32_t s_cutoff = (d == -1) ? i - 1 : n - i - 1;
  int64_t divisor = 2;
  for (uint32_t t = ceil_div(static_cast<uint32_t>(l), 2); divisor <= max_divisor;
       divisor <<= 1, t = ceil_div(static_cast<uint32_t>(l), static_cast<uint32_t>(divisor))) {
    const int32_t st = s + static_cast<int32_t>(t);
    if (st <= s_cutoff && delta(code_i, codes[i + st * d]) > delta_node) s = st;
  }
  return i + s * d + min(d, 0);
}

__global__ __launch_bounds__(256) void radix_tree_kernel(
    const uint32_t* __restrict__ codes_all, uint8_t* __restrict__ prefix_all,
    uint8_t* __restrict__ hl_all, uint8_t* __restrict__ hr_all,
    int32_t* __restrict__ lc_all, int32_t* __restrict__ par_all, SegView sv,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ n_unique,
    int32_t n_brt_host) {
  __shared__ uint32_t s_codes[256 + 2 * kHalo];
  const TileRef r = find_tile(tile_start, sv.n_seg, blockIdx.x);
  const int32_t n = n_unique ? n_unique[r.seg] - 1 : n_brt_host;
  const int32_t i0 = r.local * 256;
  if (i0 >= n) return;
  const int32_t sb = seg_begin(sv, r.seg);
  int32_t* parent = par_all + sb;
  // Stage codes [i0 - halo, i0 + 256 + halo) of the n+1 codes.
  const int32_t lo = max(0, i0 - kHalo);
  const int32_t hi = min(n + 1, i0 + 256 + kHalo);
  for (int32_t q = lo + static_cast<int32_t>(threadIdx.x); q < hi; q += 256)
    s_codes[q - lo] = codes_all[sb + q];
  __syncthreads();
  const Codes codes{codes_all + sb, s_codes, lo, hi - lo};
  const int32_t i = i0 + static_cast<int32_t>(threadIdx.x);
  if (i >= n) return;

  const uint32_t code_i = codes[i];
  int32_t d;
  int32_t l = 0;
  if (i == 0) {
    d = 1;
    l = n - 1;
  } else {
    const int32_t diff = static_cast<int32_t>(delta(code_i, codes[i + 1])) -
                         static_cast<int32_t>(delta(code_i, codes[i - 1]));
    d = (diff > 0) - (diff < 0);
    const uint32_t delta_min = delta(code_i, codes[i - d]);
    int64_t l_max = 2;
    while (i + l_max * d >= 0 && i + l_max * d <= n &&
           delta(code_i, codes[i + l_max * d]) > delta_min &&
           l_max < (int64_t(1) << 31))
      l_max *= 2;
    const int32_t l_cutoff = (d == -1) ? i : n - i;
    for (int64_t t = l_max / 2, divisor = 2; t >= 1; divisor *= 2, t = l_max / divisor) {
      if (l + t <= l_cutoff && delta(code_i, codes[i + (l + t) * d]) > delta_min)
        l += static_cast<int32_t>(t);
    }
  }
  const int32_t j = i + l * d;
  const uint32_t delta_node = delta(code_i, codes[j]);
  const int32_t gamma = split_search(codes, i, d, l, n, delta_node);
  const bool leaf_l = min(i, j) == gamma;
  const bool leaf_r = max(i, j) == gamma + 1;
  prefix_all[sb + i] = static_cast<uint8_t>(delta_node);
  hl_all[sb + i] = leaf_l;
  hr_all[sb + i] = leaf_r;
  lc_all[sb + i] = gamma;
  if (i > 0) {
    if (!leaf_l) parent[gamma] = i;
    if (!leaf_r) parent[gamma + 1] = i;
  } else {
    // The true root spans [0, n]: its internal children have no other
    // writer, and sequentially they end up holding node 0.
    parent[0] = 0;
    const uint32_t dn = delta(code_i, codes[n]);
    const int32_t g = split_search(codes, 0, 1, n, n + 1, dn);
    if (g != 0 && g < n) parent[g] = 0;
    if (g + 1 != n && g + 1 < n) parent[g + 1] = 0;
  }
}

}  // namespace

void launch_radix_tree(const uint32_t* codes, uint8_t* prefix_n,
                       uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                       int32_t* left_child, int32_t* parent, const SegView& sv,
                       const int32_t* tile_start, int32_t total_tiles,
                       const int32_t* n_unique, int32_t n_brt_host,
                       hipStream_t s) {
  if (total_tiles <= 0) return;
  radix_tree_kernel<<<total_tiles, 256, 0, s>>>(
      codes, prefix_n, has_leaf_left, has_leaf_right, left_child, parent, sv,
      tile_start, n_unique, n_brt_host);
}

}  // namespace ocp

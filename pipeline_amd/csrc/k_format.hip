// Octree output formats and a structural validator, on the device.
//
// Validator (in the role of checkTree, apps/cpu_demo/main.cpp:13-35): the
// reference's offsets are an inclusive scan used as first slots
// (kernels/06_partial_sum.cpp:3-15, include/shared/oct.h:24-90), so chains
// overlap and overwrite each other's geometry and child-in-octant does not
// hold in its own output. What holds whatever the write order, and is counted
// here per kind of violation:
//  - the written slots are slot 0 and the chain ranges [P[i], P[i] + e[i]) of
//    the radix-tree nodes i >= 1, and every chain ends inside its segment;
//    every other slot is all zero;
//  - a written slot has w = 1, a cell size range / 2^k (k = 0..11) and a
//    corner on the grid of octree level root_level + k (morton.h:45-80 with
//    oct.h:56-57, 87-88);
//  - child_leaf_mask is 0 (SetLeaf clears bits of a zeroed word, oct.h:17-21);
//  - the root slot holds the root corner and cell size = range.
// The node-mask links whose child slot lies inside the parent's octant are
// counted as information only.
//
// Compact nodes: a 12-byte header per slot - the corner as the Morton prefix
// whose decode it is, the octree level and the depth below the root (the
// cell size is range / 2^depth), the two masks, a bit per nonzero child and
// the index of its first stored child - and the nonzero children in one
// array. Expanding gives back the 60-byte nodes bit for bit. A slot that no
// chain writes has level byte 0xff.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

constexpr int kFmtThreads = 256;
constexpr int kFmtGrid = 4096;

__device__ __forceinline__ uint32_t level_shift(int32_t level) {
  return level >= 10 ? 0u : static_cast<uint32_t>(kMortonBits - 3 * level);
}

// The segment holding global slot g: the last k with oct_start[k] <= g
// (empty segments before it share its start).
__device__ __forceinline__ int32_t seg_of_slot(int64_t g, const int32_t* oct_start,
                                               int32_t n_seg) {
  int32_t lo = 0, hi = n_seg;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (oct_start[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

struct Boxes {
  const float* mn;  // per segment, or nullptr: mn1 / rg1
  const float* rg;
  float mn1, rg1;
  __device__ __forceinline__ float min_of(int32_t s) const { return mn ? mn[s] : mn1; }
  __device__ __forceinline__ float range_of(int32_t s) const { return rg ? rg[s] : rg1; }
};

// Depth k (0..11) with range / 2^k == cell bit for bit (the octree's own
// expression), or -1.
__device__ __forceinline__ int32_t depth_of(float cell, float rg) {
  for (int32_t k = 0; k <= 11; ++k) {
    const uint32_t sh = static_cast<uint32_t>(k) & 31u;
    if (__float_as_uint(rg / static_cast<float>(static_cast<int32_t>(1u << sh))) ==
        __float_as_uint(cell))
      return k;
  }
  return -1;
}

// The axis value q in [0, 1024), a multiple of 2^(10 - level), whose decode
// is c bit for bit, or -1. q is near (c - min) / range * 1024; the decode
// itself decides.
__device__ __forceinline__ int32_t axis_value(float c, int axis, int32_t level, float mn,
                                              float rg) {
  const double qd = (static_cast<double>(c) - mn) / rg * 1024.0;
  if (!(qd > -4.0 && qd < 1028.0)) return -1;
  const int32_t q0 = static_cast<int32_t>(floor(qd));
  const uint32_t grid = (1u << (level_shift(level) / 3u)) - 1u;
  for (int32_t d = -1; d <= 2; ++d) {
    const int32_t q = q0 + d;
    if (q < 0 || q > 1023 || (static_cast<uint32_t>(q) & grid)) continue;
    const uint32_t code = split_by_3(static_cast<uint32_t>(q)) << axis;
    if (__float_as_uint(decode_axis(code, axis, mn, rg)) == __float_as_uint(c)) return q;
  }
  return -1;
}

// The Morton prefix of `level` whose decode is the corner, or ~0u.
__device__ __forceinline__ uint32_t prefix_of(const float* c, int32_t level, float mn, float rg) {
  uint32_t code = 0;
  for (int a = 0; a < 3; ++a) {
    const int32_t q = axis_value(c[a], a, level, mn, rg);
    if (q < 0) return ~0u;
    code |= split_by_3(static_cast<uint32_t>(q)) << a;
  }
  return code;
}

__device__ __forceinline__ bool any_word(const OctNode& nd) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(&nd);
  uint32_t o = 0;
#pragma unroll
  for (int q = 0; q < 15; ++q) o |= w[q];
  return o != 0;
}

// Root level (prefix_n[0] / 3, 0 below two keys) and first code per segment.
__global__ void seg_roots_kernel(const uint32_t* __restrict__ codes,
                                 const uint8_t* __restrict__ prefix_n,
                                 const int32_t* __restrict__ n_unique,
                                 const int32_t* __restrict__ seg_start, int32_t n_seg,
                                 int32_t* __restrict__ root_level, uint32_t* __restrict__ root_code) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const int32_t sb = seg_start ? seg_start[s] : 0;
  const int32_t u = n_unique[s];
  root_level[s] = u >= 2 ? prefix_n[sb] / 3 : 0;
  root_code[s] = u >= 1 ? codes[sb] : 0u;
}

enum Report : int {
  kRepSlots = 0, kRepWritten, kRepCoveredUnwritten, kRepWrittenUncovered, kRepChainPastEnd,
  kRepBadW, kRepBadCell, kRepOffGrid, kRepLeafMask, kRepBadRoot, kRepLinks,
  kRepLinksInOctant, kRepFields
};

__device__ __forceinline__ void add_report(unsigned long long* rep, int field, uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, 64));
  if (lane_id() == 0 && v) atomicAdd(&rep[field], static_cast<unsigned long long>(v));
}

// cov[slot] = 1 for every slot of a chain (cov zeroed beforehand; chains of
// one segment may overlap, all store the same value). Grid: x over radix-tree
// nodes, y over segments.
__global__ __launch_bounds__(kFmtThreads) void cover_kernel(
    const int32_t* __restrict__ edge, const int32_t* __restrict__ offs,
    const int32_t* __restrict__ n_unique, const int32_t* __restrict__ oct_start,
    const int32_t* __restrict__ seg_start, int32_t n_seg, uint8_t* __restrict__ cov,
    unsigned long long* __restrict__ rep) {
  uint32_t past = 0;
  for (int32_t seg = blockIdx.y; seg < n_seg; seg += gridDim.y) {
    const int32_t m = n_unique[seg] - 1;
    const int32_t sb = seg_start ? seg_start[seg] : 0;
    const int32_t nb = oct_start[seg];
    const int32_t k = oct_start[seg + 1] - nb;
    for (int32_t i = blockIdx.x * kFmtThreads + threadIdx.x + 1; i < m;
         i += gridDim.x * kFmtThreads) {
      const int32_t e = edge[sb + i];
      if (e <= 0) continue;
      const int32_t p = offs[sb + i];
      if (p < 0 || p > k - e) {
        ++past;
        continue;
      }
      for (int32_t j = 0; j < e; ++j) cov[nb + p + j] = 1;
    }
  }
  add_report(rep, kRepChainPastEnd, past);
}

__global__ __launch_bounds__(kFmtThreads) void validate_kernel(
    const OctNode* __restrict__ nodes, const uint8_t* __restrict__ cov,
    const int32_t* __restrict__ oct_start, int32_t n_seg, Boxes bx,
    const int32_t* __restrict__ root_level, const uint32_t* __restrict__ root_code,
    unsigned long long* __restrict__ rep) {
  uint32_t cnt[kRepFields] = {};
  const int64_t k_total = oct_start[n_seg];
  for (int64_t g = blockIdx.x * int64_t(kFmtThreads) + threadIdx.x; g < k_total;
       g += int64_t(gridDim.x) * kFmtThreads) {
    const int32_t seg = seg_of_slot(g, oct_start, n_seg);
    const int32_t base = oct_start[seg];
    const int32_t local = static_cast<int32_t>(g - base);
    const int32_t k = oct_start[seg + 1] - base;
    const OctNode nd = nodes[g];
    const bool written = any_word(nd);
    const bool covered = local == 0 || cov[g] != 0;
    ++cnt[kRepSlots];
    cnt[kRepWritten] += written;
    cnt[kRepCoveredUnwritten] += covered && !written;
    cnt[kRepWrittenUncovered] += written && !covered;
    if (!written) continue;
    const float mn = bx.min_of(seg), rg = bx.range_of(seg);
    const int32_t rl = root_level[seg];
    cnt[kRepBadW] += __float_as_uint(nd.corner[3]) != __float_as_uint(1.0f);
    cnt[kRepLeafMask] += nd.child_leaf_mask != 0;
    const int32_t depth = depth_of(nd.cell_size, rg);
    if (depth < 0) {
      ++cnt[kRepBadCell];
      continue;
    }
    const uint32_t pre = prefix_of(nd.corner, rl + depth, mn, rg);
    cnt[kRepOffGrid] += pre == ~0u;
    if (local == 0) {
      const uint32_t sh = level_shift(rl);
      const uint32_t c0 = (root_code[seg] >> sh) << sh;
      bool same = depth == 0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        same &= __float_as_uint(decode_axis(c0, a, mn, rg)) == __float_as_uint(nd.corner[a]);
      cnt[kRepBadRoot] += !same;
    }
    const float h = nd.cell_size * 0.5f;
    for (int c = 0; c < 8; ++c) {
      if (!((nd.child_node_mask >> c) & 1)) continue;
      ++cnt[kRepLinks];
      const int32_t t = nd.children[c];
      if (t < 0 || t >= k) continue;
      const OctNode ch = nodes[base + t];
      const float lo0 = nd.corner[0] + ((c & 1) ? h : 0.0f);
      const float lo1 = nd.corner[1] + ((c & 2) ? h : 0.0f);
      const float lo2 = nd.corner[2] + ((c & 4) ? h : 0.0f);
      cnt[kRepLinksInOctant] += ch.cell_size > 0.0f && ch.cell_size <= h &&
                                ch.corner[0] >= lo0 && ch.corner[1] >= lo1 &&
                                ch.corner[2] >= lo2 && ch.corner[0] + ch.cell_size <= lo0 + h &&
                                ch.corner[1] + ch.cell_size <= lo1 + h &&
                                ch.corner[2] + ch.cell_size <= lo2 + h;
    }
  }
#pragma unroll
  for (int f = 0; f < kRepFields; ++f)
    if (f != kRepChainPastEnd) add_report(rep, f, cnt[f]);
}

// ---- compact form
struct CompactNode {
  uint32_t prefix;
  uint8_t level;  // level | depth << 4; 0xff: never written
  uint8_t node_mask;
  uint8_t child_set;
  uint8_t leaf_mask;
  uint32_t first_child;
};
static_assert(sizeof(CompactNode) == 12, "compact node layout");

// The header of a slot without its first-child index, or false when the
// slot has no exact compact form.
__device__ __forceinline__ bool compact_header(const OctNode& nd, float mn, float rg, int32_t rl,
                                               CompactNode* cn) {
  *cn = CompactNode{};
  uint32_t set = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) set |= (nd.children[q] != 0 ? 1u : 0u) << q;
  cn->child_set = static_cast<uint8_t>(set);
  if (!any_word(nd)) {
    cn->level = 0xff;
    return true;
  }
  const int32_t depth = depth_of(nd.cell_size, rg);
  if (depth < 0 || rl + depth > 14) return false;
  const uint32_t pre = prefix_of(nd.corner, rl + depth, mn, rg);
  if (pre == ~0u || __float_as_uint(nd.corner[3]) != __float_as_uint(1.0f) ||
      static_cast<uint32_t>(nd.child_node_mask) > 255u ||
      static_cast<uint32_t>(nd.child_leaf_mask) > 255u)
    return false;
  cn->prefix = pre;
  cn->level = static_cast<uint8_t>((rl + depth) | (depth << 4));
  cn->node_mask = static_cast<uint8_t>(nd.child_node_mask);
  cn->leaf_mask = static_cast<uint8_t>(nd.child_leaf_mask);
  return true;
}

// Pass 1: nonzero children per slot into cnt, and a count of slots with no
// exact compact form.
__global__ __launch_bounds__(kFmtThreads) void compact_count_kernel(
    const OctNode* __restrict__ nodes, const int32_t* __restrict__ oct_start, int32_t n_seg,
    Boxes bx, const int32_t* __restrict__ root_level, int32_t* __restrict__ cnt,
    unsigned long long* __restrict__ bad) {
  const int64_t k_total = oct_start[n_seg];
  uint32_t nbad = 0;
  for (int64_t g = blockIdx.x * int64_t(kFmtThreads) + threadIdx.x; g < k_total;
       g += int64_t(gridDim.x) * kFmtThreads) {
    const OctNode nd = nodes[g];
    const int32_t seg = seg_of_slot(g, oct_start, n_seg);
    CompactNode cn;
    nbad += !compact_header(nd, bx.min_of(seg), bx.range_of(seg), root_level[seg], &cn);
    cnt[g] = __popc(cn.child_set);
  }
  add_report(bad, 0, nbad);
}

// Pass 2: headers and children; incl is the inclusive scan of the counts.
__global__ __launch_bounds__(kFmtThreads) void compact_write_kernel(
    const OctNode* __restrict__ nodes, const int32_t* __restrict__ oct_start, int32_t n_seg,
    Boxes bx, const int32_t* __restrict__ root_level, const int32_t* __restrict__ incl,
    CompactNode* __restrict__ out, int32_t* __restrict__ children) {
  const int64_t k_total = oct_start[n_seg];
  for (int64_t g = blockIdx.x * int64_t(kFmtThreads) + threadIdx.x; g < k_total;
       g += int64_t(gridDim.x) * kFmtThreads) {
    const OctNode nd = nodes[g];
    const int32_t seg = seg_of_slot(g, oct_start, n_seg);
    CompactNode cn;
    (void)compact_header(nd, bx.min_of(seg), bx.range_of(seg), root_level[seg], &cn);
    int32_t o = incl[g] - __popc(cn.child_set);
    cn.first_child = static_cast<uint32_t>(o);
    for (int q = 0; q < 8; ++q)
      if (nd.children[q] != 0) children[o++] = nd.children[q];
    out[g] = cn;
  }
}

// 60-byte nodes from the compact form. Child indices at or past n_children
// read as 0.
__global__ __launch_bounds__(kFmtThreads) void expand_kernel(
    const CompactNode* __restrict__ in, const int32_t* __restrict__ children, int64_t n_children,
    const int32_t* __restrict__ oct_start, int32_t n_seg, Boxes bx, OctNode* __restrict__ out) {
  const int64_t k_total = oct_start[n_seg];
  for (int64_t g = blockIdx.x * int64_t(kFmtThreads) + threadIdx.x; g < k_total;
       g += int64_t(gridDim.x) * kFmtThreads) {
    const CompactNode cn = in[g];
    OctNode nd{};
    int64_t o = cn.first_child;
    for (int q = 0; q < 8; ++q) {
      if (!((cn.child_set >> q) & 1u)) continue;
      // a malformed node (offset outside [0, n_children)) reads no memory
      nd.children[q] = (o >= 0 && o < n_children) ? children[o] : 0;
      ++o;
    }
    if (cn.level != 0xff) {
      const int32_t seg = seg_of_slot(g, oct_start, n_seg);
      const float mn = bx.min_of(seg), rg = bx.range_of(seg);
      nd.corner[0] = decode_axis(cn.prefix, 0, mn, rg);
      nd.corner[1] = decode_axis(cn.prefix, 1, mn, rg);
      nd.corner[2] = decode_axis(cn.prefix, 2, mn, rg);
      nd.corner[3] = 1.0f;
      const uint32_t sh = static_cast<uint32_t>(cn.level >> 4) & 31u;
      nd.cell_size = rg / static_cast<float>(static_cast<int32_t>(1u << sh));
      nd.child_node_mask = cn.node_mask;
      nd.child_leaf_mask = cn.leaf_mask;
    }
    out[g] = nd;
  }
}

// ---- inclusive scan of n int32 counts in place (three launches, no
// inter-workgroup waits). bsum: n_blocks + 1 words; bsum[n_blocks] = total.
constexpr int kScanItems = 16;
constexpr int kScanBlock = kFmtThreads * kScanItems;

__global__ __launch_bounds__(kFmtThreads) void scan_reduce_kernel(const int32_t* __restrict__ a,
                                                                  int64_t n,
                                                                  uint32_t* __restrict__ bsum) {
  __shared__ uint32_t tmp[kFmtThreads / 64];
  const int64_t b0 = int64_t(blockIdx.x) * kScanBlock;
  uint32_t s = 0;
  for (int u = 0; u < kScanItems; ++u) {
    const int64_t i = b0 + int64_t(u) * kFmtThreads + threadIdx.x;
    if (i < n) s += static_cast<uint32_t>(a[i]);
  }
  uint32_t total;
  (void)block_excl_scan<kFmtThreads>(s, tmp, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t* __restrict__ bsum,
                                                           int32_t nb) {
  __shared__ uint32_t tmp[16];
  uint32_t carry = 0;
  for (int32_t c = 0; c < nb; c += 1024) {
    const int32_t k = c + static_cast<int32_t>(threadIdx.x);
    const uint32_t v = k < nb ? bsum[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan<1024>(v, tmp, &total);
    if (k < nb) bsum[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kFmtThreads) void scan_apply_kernel(int32_t* __restrict__ a,
                                                                 int64_t n,
                                                                 const uint32_t* __restrict__ bsum) {
  __shared__ uint32_t tmp[kFmtThreads / 64];
  const int64_t t0 = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int u = 0; u < kScanItems; ++u) {
    v[u] = t0 + u < n ? static_cast<uint32_t>(a[t0 + u]) : 0u;
    s += v[u];
  }
  uint32_t total;
  uint32_t run = bsum[blockIdx.x] + block_excl_scan<kFmtThreads>(s, tmp, &total);
#pragma unroll
  for (int u = 0; u < kScanItems; ++u) {
    run += v[u];
    if (t0 + u < n) a[t0 + u] = static_cast<int32_t>(run);
  }
}

}  // namespace

void launch_seg_roots(const uint32_t* codes, const uint8_t* prefix_n, const int32_t* n_unique,
                      const int32_t* seg_start, int32_t n_seg, int32_t* root_level,
                      uint32_t* root_code, hipStream_t s) {
  if (n_seg <= 0) return;
  seg_roots_kernel<<<(n_seg + 255) / 256, 256, 0, s>>>(codes, prefix_n, n_unique, seg_start,
                                                       n_seg, root_level, root_code);
}

void launch_validate(const OctNode* nodes, const int32_t* edge, const int32_t* offs,
                     const int32_t* n_unique, const int32_t* oct_start, const int32_t* seg_start,
                     int32_t n_seg, const float* smin, const float* srg, float mn1, float rg1,
                     const int32_t* root_level, const uint32_t* root_code, uint8_t* cov,
                     unsigned long long* rep, hipStream_t s) {
  if (n_seg <= 0) return;
  dim3 grid(64, std::min(n_seg, 1024));
  cover_kernel<<<grid, kFmtThreads, 0, s>>>(edge, offs, n_unique, oct_start, seg_start, n_seg,
                                            cov, rep);
  const Boxes bx{smin, srg, mn1, rg1};
  validate_kernel<<<kFmtGrid, kFmtThreads, 0, s>>>(nodes, cov, oct_start, n_seg, bx, root_level,
                                                   root_code, rep);
}

int64_t compact_scan_blocks(int64_t k_total) { return (k_total + kScanBlock - 1) / kScanBlock; }

void launch_compact_count(const OctNode* nodes, const int32_t* oct_start, int32_t n_seg,
                          const float* smin, const float* srg, float mn1, float rg1,
                          const int32_t* root_level, int64_t k_total, int32_t* cnt,
                          uint32_t* bsum, unsigned long long* bad, hipStream_t s) {
  if (n_seg <= 0) return;
  const Boxes bx{smin, srg, mn1, rg1};
  compact_count_kernel<<<kFmtGrid, kFmtThreads, 0, s>>>(nodes, oct_start, n_seg, bx, root_level,
                                                        cnt, bad);
  const int64_t nb = compact_scan_blocks(k_total);
  if (nb > 0)
    scan_reduce_kernel<<<static_cast<uint32_t>(nb), kFmtThreads, 0, s>>>(cnt, k_total, bsum);
  scan_blocks_kernel<<<1, 1024, 0, s>>>(bsum, static_cast<int32_t>(nb));
  if (nb > 0)
    scan_apply_kernel<<<static_cast<uint32_t>(nb), kFmtThreads, 0, s>>>(cnt, k_total, bsum);
}

void launch_compact_write(const OctNode* nodes, const int32_t* oct_start, int32_t n_seg,
                          const float* smin, const float* srg, float mn1, float rg1,
                          const int32_t* root_level, const int32_t* incl, void* out,
                          int32_t* children, hipStream_t s) {
  if (n_seg <= 0) return;
  const Boxes bx{smin, srg, mn1, rg1};
  compact_write_kernel<<<kFmtGrid, kFmtThreads, 0, s>>>(nodes, oct_start, n_seg, bx, root_level,
                                                        incl, static_cast<CompactNode*>(out),
                                                        children);
}

void launch_expand(const void* in, const int32_t* children, int64_t n_children,
                   const int32_t* oct_start, int32_t n_seg, const float* smin, const float* srg,
                   float mn1, float rg1, OctNode* out, hipStream_t s) {
  if (n_seg <= 0) return;
  const Boxes bx{smin, srg, mn1, rg1};
  expand_kernel<<<kFmtGrid, kFmtThreads, 0, s>>>(static_cast<const CompactNode*>(in), children,
                                                 n_children, oct_start, n_seg, bx, out);
}

}  // namespace ocp

// Stage 7 for the per-stage entry points: k_MakeOctNodes and
// k_LinkLeafNodes called separately with the reference's argument lists
// (include/cuda/kernels/07_octree.cuh:7-28). The make call receives only
// parent links, not the children, so its top links are pushed upwards here
// instead of gathered downwards as in the fused build (k_octree.hip).
// Same sequential (ascending index) semantics of include/shared/oct.h:24-134.
//
// scatter (one thread per BRT node): every child-slot write of the
//   sequential program is an atomicMax of an order key into the slot's
//   children[c] word, used as a 32-bit unsigned scratch: make writes carry
//   (i << 4 | j) + 1 (chain i, step j), link writes 0x80000000 | (i << 1 |
//   side), so the maximum is the last writer and every link write beats
//   every make write, as the link pass runs after the make pass. Child-node
//   mask bits are atomicOr (OR commutes).
// finalize (one thread per BRT node): chain i owns slot P[i]+j when no
//   later chain covers it (later non-empty chains start strictly later, so
//   a short forward walk over P finds them). The owner writes corner and
//   cell size from (i, j) and turns the children keys of its slots back
//   into values: P[i'] + j' for a make key, left_child[i'] + side for a
//   link key. Every slot that receives a child write is either covered by
//   a chain or is the root slot 0, which thread 0 finalises.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

struct SegBind {
  OctNode* nodes;
  int32_t k;
  int32_t m;
  const uint32_t* codes;
  const uint8_t* pn;
  const uint8_t* hl;
  const uint8_t* hr;
  const int32_t* lc;
  const int32_t* par;
  const int32_t* e;
  const int32_t* p;
  float mn;
  float rg;
};

// Binds thread -> (segment, BRT node i). With root_only, thread 0 of a
// segment that holds a single unique key (no BRT node, one root slot) is
// bound too, with m = 0.
__device__ __forceinline__ bool bind(const OctArgs& a, const SegView& sv,
                                     const int32_t* tile_start, int32_t& i,
                                     SegBind& b, bool root_only = false) {
  if (a.ctrl && a.ctrl[kCtrlOctOk] == 0) return false;
  const TileRef r = find_tile(tile_start, sv.n_seg, blockIdx.x);
  b.m = a.n_unique ? a.n_unique[r.seg] - 1 : a.n_brt_host;
  i = r.local * kNodeTile + static_cast<int32_t>(threadIdx.x);
  if (i >= b.m && !(root_only && i == 0 && b.m == 0)) return false;
  const int32_t sb = seg_begin(sv, r.seg);
  b.nodes = a.nodes + (a.oct_start ? a.oct_start[r.seg] : 0);
  b.k = a.kalloc ? a.kalloc[r.seg] : a.k_host;
  b.codes = a.codes + sb;
  b.pn = a.prefix_n + sb;
  b.hl = a.has_leaf_left + sb;
  b.hr = a.has_leaf_right + sb;
  b.lc = a.left_child + sb;
  b.par = a.parent + sb;
  b.e = a.edge + sb;
  b.p = a.offsets + sb;
  b.mn = seg_min(sv, r.seg);
  b.rg = seg_range(sv, r.seg);
  return true;
}

__device__ __forceinline__ uint32_t child_of(uint32_t code, int32_t level) {
  return (code >> (static_cast<uint32_t>(kMortonBits - 3 * level) & 31u)) & 7u;
}

__device__ __forceinline__ void put_key(const SegBind& b, int32_t slot,
                                        uint32_t which, uint32_t key,
                                        bool node_bit) {
  if (slot < 0 || slot >= b.k) return;
  OctNode* nd = b.nodes + slot;
  atomicMax(reinterpret_cast<uint32_t*>(&nd->children[which]), key);
  if (node_bit) atomicOr(&nd->child_node_mask, 1 << which);
}

__global__ __launch_bounds__(kNodeTile) void oct_scatter_kernel(
    OctArgs a, SegView sv, const int32_t* __restrict__ tile_start, int phases) {
  int32_t i;
  SegBind b;
  if (!bind(a, sv, tile_start, i, b)) return;
  if ((phases & 1) && i >= 1) {
    const int32_t n_new = b.e[i];
    if (n_new > 0) {
      const int32_t p_i = b.p[i];
      const int32_t lvl = b.pn[i] / 3;
      const uint32_t code = b.codes[i];
      for (int32_t j = 0; j < n_new - 1; ++j)
        put_key(b, p_i + j + 1, child_of(code, lvl - j),
                static_cast<uint32_t>((i << 4) | j) + 1u, true);
      // oct.h:61-73: first ancestor with edges, at most 31 hops.
      int32_t rt = b.par[i];
      int32_t counter = 0;
      while (b.e[rt] == 0) {
        rt = b.par[rt];
        if (++counter > 30) break;
      }
      put_key(b, b.p[rt], child_of(code, lvl - n_new + 1),
              static_cast<uint32_t>((i << 4) | (n_new - 1)) + 1u, true);
    }
  }
  if (phases & 2) {
    const int32_t leaf_level = b.pn[i] / 3 + 1;
    int32_t anc = -1;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const bool has = side == 0 ? b.hl[i] : b.hr[i];
      if (!has) continue;
      const int32_t leaf = b.lc[i] + side;
      const uint32_t which = child_of(b.codes[leaf], leaf_level);
      if (anc < 0) {
        // oct.h:111-113 stopped at the root and after 64 hops.
        anc = i;
        int32_t hops = 0;
        while (b.e[anc] == 0 && anc != 0 && hops < 64) {
          anc = b.par[anc];
          ++hops;
        }
      }
      const int32_t slot = b.p[anc];
      put_key(b, slot, which, 0x80000000u | static_cast<uint32_t>((i << 1) | side), false);
      if ((phases & 4) && slot >= 0 && slot < b.k)
        atomicAnd(&b.nodes[slot].child_leaf_mask, ~(1 << which));
    }
  }
}

__device__ __forceinline__ void resolve_children(const SegBind& b, int32_t slot,
                                                 int resolve) {
  int32_t* ch = b.nodes[slot].children;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t v = static_cast<uint32_t>(ch[c]);
    if (v == 0) continue;
    if (v & 0x80000000u) {
      if (resolve & 2) {
        const uint32_t q = v & 0x7fffffffu;
        ch[c] = b.lc[q >> 1] + static_cast<int32_t>(q & 1u);
      }
    } else if (resolve & 1) {
      const uint32_t q = v - 1u;
      ch[c] = b.p[q >> 4] + static_cast<int32_t>(q & 15u);
    }
  }
}

__device__ __forceinline__ void write_geometry(const SegBind& b, int32_t slot,
                                               uint32_t code, int32_t level,
                                               int32_t root_level) {
  const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * level) & 31u;
  const uint32_t c = (code >> shift) << shift;
  OctNode* nd = b.nodes + slot;
  nd->corner[0] = decode_axis(c, 0, b.mn, b.rg);
  nd->corner[1] = decode_axis(c, 1, b.mn, b.rg);
  nd->corner[2] = decode_axis(c, 2, b.mn, b.rg);
  nd->corner[3] = 1.0f;
  const uint32_t sh = static_cast<uint32_t>(level - root_level) & 31u;
  nd->cell_size = b.rg / static_cast<float>(static_cast<int32_t>(1u << sh));
}

__global__ __launch_bounds__(kNodeTile) void oct_finalize_kernel(
    OctArgs a, SegView sv, const int32_t* __restrict__ tile_start,
    int resolve) {
  int32_t i;
  SegBind b;
  if (!bind(a, sv, tile_start, i, b, true)) return;
  const int32_t root_level = b.m >= 1 ? b.pn[0] / 3 : 0;
  if (i == 0) {
    if (b.k < 1) return;
    if (resolve & 8) {
      // Root: corner of codes[0]'s root-level prefix, cell size = range.
      const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * root_level) & 31u;
      const uint32_t c = (b.codes[0] >> shift) << shift;
      OctNode* nd = b.nodes;
      nd->corner[0] = decode_axis(c, 0, b.mn, b.rg);
      nd->corner[1] = decode_axis(c, 1, b.mn, b.rg);
      nd->corner[2] = decode_axis(c, 2, b.mn, b.rg);
      nd->corner[3] = 1.0f;
      nd->cell_size = b.rg;
    }
    resolve_children(b, 0, resolve);
    return;
  }
  const int32_t n_new = b.e[i];
  if (n_new <= 0) return;
  const int32_t p_i = b.p[i];
  const int32_t last = p_i + n_new - 1;
  uint32_t owned = (n_new >= 32) ? ~0u : ((1u << n_new) - 1u);
  for (int32_t k = i + 1; k < b.m && owned; ++k) {
    const int32_t pk = b.p[k];
    if (pk > last) break;
    const int32_t ek = b.e[k];
    if (ek <= 0) continue;
    const int32_t lo = max(pk, p_i) - p_i;
    const int32_t hi = min(pk + ek - 1, last) - p_i;
    for (int32_t q = lo; q <= hi; ++q) owned &= ~(1u << q);
  }
  const int32_t lvl = b.pn[i] / 3;
  const uint32_t code = b.codes[i];
  for (int32_t j = 0; j < n_new && j < 32; ++j) {
    if (!((owned >> j) & 1u)) continue;
    const int32_t slot = p_i + j;
    // Slot 0 belongs to the root thread (chains start at P[i] >= e[i] > 0
    // for any tree the earlier stages produce).
    if (slot <= 0 || slot >= b.k) continue;
    if (resolve & 4) write_geometry(b, slot, code, lvl - j, root_level);
    resolve_children(b, slot, resolve);
  }
}

}  // namespace

void launch_oct_scatter(const OctArgs& a, const SegView& sv,
                        const int32_t* tile_start, int32_t total_tiles,
                        int phases, hipStream_t s) {
  if (total_tiles <= 0) return;
  oct_scatter_kernel<<<total_tiles, kNodeTile, 0, s>>>(a, sv, tile_start, phases);
}

void launch_oct_finalize(const OctArgs& a, const SegView& sv,
                         const int32_t* tile_start, int32_t total_tiles,
                         int resolve, hipStream_t s) {
  if (total_tiles <= 0) return;
  oct_finalize_kernel<<<total_tiles, kNodeTile, 0, s>>>(a, sv, tile_start, resolve);
}

}  // namespace ocp

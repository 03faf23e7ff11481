// Stage 7 of the fused build: octree nodes from the radix tree -
// ProcessOctNode and ProcessLinkLeaf of include/shared/oct.h:24-134, driven
// by kernels/07_octree.cpp:5-57 - with the sequential (ascending index)
// semantics of the reference CPU path, in one pass and without atomics.
//
// The reference's chains overlap: the offsets P are an inclusive scan
// (kernels/06_partial_sum.cpp) but P[i] is used as the first slot of chain
// i, so run sequentially the last writer of a field wins. The kernel gives
// every node slot exactly one writer that computes its final 60 bytes:
//
//  * Chain i (e[i] > 0) covers slots P[i] .. P[i]+e[i]-1. Non-empty chains
//    start strictly later as i grows, so chain i always owns its first slot
//    and owns slot P[i]+j unless a later chain covers it. The owner writes
//    the corner and cell size of step j (oct.h:44-57, 76-88).
//  * Make-phase child links into slot s come from the chains covering s:
//    chain q writes children[c] = s-1 there (oct.h:41-51), so those links
//    need no ordering, only the union of their child indices. The first
//    slot P[a] of chain a additionally receives the top links of the
//    non-empty chains whose parent walk (oct.h:61-73) ends at a, and the
//    leaf links (oct.h:92-134) of every node whose walk ends at a. Those
//    nodes are a's group: a and the e == 0 internal nodes below it, at most
//    two levels deep because prefix lengths grow strictly down the tree and
//    one octree level spans three of them. The owner visits the group
//    downwards and keeps, per child index, the write that comes last in
//    the sequential order: make writes by (chain, step), every link write
//    after every make write, links by (node, side).
//  * The root group follows parent[] rather than node 0's own children:
//    node 0 is the reference's short-range root (04_radix_tree.cpp:63-65),
//    and the nodes whose parent is 0 are the children of the true root.
//  * Slots no chain covers stay zero; the chain before the gap (or the root
//    for the slots before the first chain) writes those zeros, so the node
//    buffer needs no zero fill.
//  * child_leaf_mask only ever sees AND-NOT on a zeroed word (oct.h:20), so
//    it is 0; child_node_mask is the OR of the make-phase link bits.
//
// Layout of the work: a workgroup of 128 threads takes 256 consecutive BRT
// nodes and stages two packed 8-byte records {P, e + prefix length} and
// {left child + leaf flags, code} of those nodes plus 32 on either side into
// LDS with coalesced loads. It compacts the non-empty chains of its range
// so that every lane processes a chain. The short walks over neighbouring
// chains and most child lookups of the group walk are LDS reads (anything
// outside the window falls back to global memory); the backward walk stops
// as soon as no earlier chain can reach this one, using the largest edge
// count of the window. The group walk is a depth-first loop per lane, so a
// wave runs as many steps as its largest group rather than a fixed tree of
// visits. Each chain keeps the order key of the last writer per child index
// of its first slot in an LDS column and decodes the values when it writes
// the node. Nodes are assembled in registers and written with three 16-byte
// and one 12-byte store.
//
// Edge counts are at most 10 (prefix lengths are at most 31 for 30-bit
// codes), so a chain spans at most 10 slots.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

constexpr int32_t kMaxEdge = 10;
constexpr int kThreads = 128;
#ifndef OCP_OCT_HALO
#define OCP_OCT_HALO 32
#endif
// BRT nodes staged on either side of a tile. With 16-byte window records
// and a 32-bit last-writer table: 10M uniform points, halo 8 -> 330 us,
// 16 -> 308, 32 -> 286, 64 -> 296, 128 -> 340, 256 -> 422 (LDS per
// workgroup bounded the waves per CU). At 8 waves per SIMD with 8-byte
// records, halos up to 48 keep that occupancy and time the same
// (32 -> 256 us, 40 -> 258, 48 -> 254).
constexpr int kHalo = OCP_OCT_HALO;
constexpr int kWinN = kNodeTile + 2 * kHalo;
constexpr uint32_t kLeafL = 1u << 30;
constexpr uint32_t kLeafR = 1u << 31;
constexpr uint32_t kChildMask = (1u << 30) - 1u;

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef int32_t v3i __attribute__((ext_vector_type(3)));
typedef int32_t v2i __attribute__((ext_vector_type(2)));

// A window node is two 8-byte LDS records: {P, e | prefix length << 16} and
// {left child | leaf flags, code}. Edge counts are small signed values (at
// most 85 in magnitude for byte prefix lengths) and prefix lengths bytes, so
// both fit one word, and 16 bytes per node (20 with a separate code array)
// keep the workgroup's LDS within what 8 waves per SIMD allow.
__device__ __forceinline__ int32_t pack_epn(int32_t e, uint32_t pn) {
  return static_cast<int32_t>((static_cast<uint32_t>(e) & 0xffffu) | (pn << 16));
}
__device__ __forceinline__ int32_t epn_e(int32_t w) {
  return static_cast<int32_t>(static_cast<int16_t>(static_cast<uint32_t>(w) & 0xffffu));
}
__device__ __forceinline__ int32_t epn_pn(int32_t w) {
  return static_cast<int32_t>(static_cast<uint32_t>(w) >> 16);
}

struct SegBind {
  OctNode* nodes;
  int32_t k;
  int32_t m;
  const uint32_t* codes;
  const uint8_t* pn;
  const uint8_t* hl;
  const uint8_t* hr;
  const int32_t* lc;
  const int32_t* e;
  const int32_t* p;
  float mn;
  float rg;
};

// One BRT node as the kernel reads it.
struct Rec {
  int32_t p;
  int32_t e;
  uint32_t lc;  // left child | kLeafL | kLeafR
  int32_t pn;
  __device__ __forceinline__ int32_t child() const { return static_cast<int32_t>(lc & kChildMask); }
  __device__ __forceinline__ bool leaf(int side) const { return (lc & (side ? kLeafR : kLeafL)) != 0; }
};

struct View {
  const SegBind& b;
  int32_t lo;       // window [lo, lo + n) of nodes, [lo, lo + nc) of codes
  int32_t n;
  int32_t nc;
  lds_ptr<const v2i> pe;   // {P, e | pn << 16}
  lds_ptr<const v2i> lcc;  // {left child | flags, code}
  lds_ptr<uint32_t> sc_col;  // this thread's SlotChildren column
  __device__ __forceinline__ Rec R(int32_t q) const {
    if (static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n)) {
      const v2i a = pe[q - lo];
      const v2i c = lcc[q - lo];
      return {a.x, epn_e(a.y), static_cast<uint32_t>(c.x), epn_pn(a.y)};
    }
    const uint32_t l = static_cast<uint32_t>(b.lc[q]) | (b.hl[q] ? kLeafL : 0u) |
                       (b.hr[q] ? kLeafR : 0u);
    return {b.p[q], b.e[q], l, static_cast<int32_t>(b.pn[q])};
  }
  __device__ __forceinline__ int32_t P(int32_t q) const {
    return static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n) ? pe[q - lo].x : b.p[q];
  }
  __device__ __forceinline__ int32_t E(int32_t q) const {
    return static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n) ? epn_e(pe[q - lo].y) : b.e[q];
  }
  __device__ __forceinline__ uint32_t CODE(int32_t q) const {
    return static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(nc)
               ? static_cast<uint32_t>(lcc[q - lo].y) : b.codes[q];
  }
  // Record and code of internal node q (q < m) under one window test: the
  // code window is at least as long as the node window.
  __device__ __forceinline__ Rec RC(int32_t q, uint32_t* c) const {
    if (static_cast<uint32_t>(q - lo) < static_cast<uint32_t>(n)) {
      const v2i a = pe[q - lo];
      const v2i l = lcc[q - lo];
      *c = static_cast<uint32_t>(l.y);
      return {a.x, epn_e(a.y), static_cast<uint32_t>(l.x), epn_pn(a.y)};
    }
    *c = b.codes[q];
    const uint32_t l = static_cast<uint32_t>(b.lc[q]) | (b.hl[q] ? kLeafL : 0u) |
                       (b.hr[q] ? kLeafR : 0u);
    return {b.p[q], b.e[q], l, static_cast<int32_t>(b.pn[q])};
  }
};

__device__ __forceinline__ uint32_t child_of(uint32_t code, int32_t level) {
  return (code >> (static_cast<uint32_t>(kMortonBits - 3 * level) & 31u)) & 7u;
}
__device__ __forceinline__ uint32_t delta(uint32_t a, uint32_t b) {
  const uint32_t x = a ^ b;
  return ((x ? static_cast<uint32_t>(__clz(static_cast<int>(x))) : 32u) - 1u) & 0xffu;
}

// Per-child-index last writer of one slot, as its order key only: a make
// write (chain q, step t) always stores P[q] + t and a link write (node x,
// side) stores left_child[x] + side, so the value follows from the key.
// Writers have distinct keys and the largest key is the last writer; key 0
// is the untouched state (value 0). One LDS word per child index in the
// thread's own column, kept by an LDS atomic max: a register table needs a
// compare-and-select over all eight entries per write.
struct SlotChildren {
  lds_ptr<uint32_t> col;  // entry c at col[c * kThreads]
  uint32_t node_mask;
  __device__ __forceinline__ void init(lds_ptr<uint32_t> column) {
    col = column;
#pragma unroll
    for (int c = 0; c < 8; ++c) col[c * kThreads] = 0;
    node_mask = 0;
  }
  __device__ __forceinline__ void put(uint32_t which, uint32_t k) {
    __hip_atomic_fetch_max(col + which * kThreads, k, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ uint32_t key(int c) const { return col[c * kThreads]; }
};

__device__ __forceinline__ uint32_t make_key(int32_t chain, int32_t step) {
  return static_cast<uint32_t>((chain << 4) | step) + 1u;
}
__device__ __forceinline__ uint32_t link_key(int32_t node, int side) {
  return 0x80000000u | static_cast<uint32_t>((node << 1) | side);
}
// The child value a write with this key stores.
__device__ __forceinline__ int32_t key_value(const View& v, uint32_t k) {
  if (k == 0) return 0;
  if (k & 0x80000000u) {
    const int32_t x = static_cast<int32_t>((k & 0x7fffffffu) >> 1);
    return v.R(x).child() + static_cast<int32_t>(k & 1u);
  }
  return v.P(static_cast<int32_t>((k - 1u) >> 4)) + static_cast<int32_t>((k - 1u) & 15u);
}

// Child `side` of group node x (record rx): a leaf link, a top link of a
// non-empty chain, or (returned) an e == 0 internal node joining the group.
__device__ __forceinline__ int32_t visit_child(const View& v, int32_t x, const Rec& rx,
                                               int side, SlotChildren& sc, Rec* out) {
  const int32_t ch = rx.child() + side;
  if (rx.leaf(side)) {
    if (ch < 0 || ch > v.b.m) return -1;  // leaves index the m+1 codes
    sc.put(child_of(v.CODE(ch), rx.pn / 3 + 1), link_key(x, side));
    return -1;
  }
  if (ch <= 0 || ch >= v.b.m) return -1;
  uint32_t code_ch;
  const Rec ry = v.RC(ch, &code_ch);
  if (ry.e == 0) {
    *out = ry;
    return ch;
  }
  if (ry.e > 0) {
    const uint32_t w = child_of(code_ch, ry.pn / 3 - ry.e + 1);
    sc.put(w, make_key(ch, ry.e - 1));
    sc.node_mask |= 1u << w;
  }
  return -1;
}

// Depth-first walk of a group from node x at depth d: both child visits of
// every member down to depth dmax (e == 0 children below that are not
// followed). Four pending members suffice: each level below x leaves at
// most one sibling behind, plus the two children of the deepest level.
__device__ __forceinline__ void walk_group(const View& v, int32_t x, Rec rx, int d, int dmax,
                                           SlotChildren& sc) {
  int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // pending members, s0 on top
  uint32_t ds = 0;                         // their depths, 4 bits each, top lowest
  int n = 0;
  while (true) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      Rec ry;
      const int32_t y = visit_child(v, x, rx, side, sc, &ry);
      if (y >= 0 && d < dmax) {
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = y;
        ds = (ds << 4) | static_cast<uint32_t>(d + 1);
        ++n;
      }
    }
    if (n == 0) break;
    x = s0;
    d = static_cast<int>(ds & 15u);
    s0 = s1;
    s1 = s2;
    s2 = s3;
    ds >>= 4;
    --n;
    rx = v.R(x);
  }
}

// A whole node in registers: children, corner, cell size, masks.
struct NodeOut {
  int32_t w[15];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int c = 0; c < 15; ++c) w[c] = 0;
  }
  __device__ __forceinline__ void geometry(uint32_t code, int32_t level, int32_t root_level,
                                           float mn, float rg) {
    const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * level) & 31u;
    const uint32_t c = (code >> shift) << shift;
    w[8] = __float_as_int(decode_axis(c, 0, mn, rg));
    w[9] = __float_as_int(decode_axis(c, 1, mn, rg));
    w[10] = __float_as_int(decode_axis(c, 2, mn, rg));
    w[11] = __float_as_int(1.0f);
    const uint32_t sh = static_cast<uint32_t>(level - root_level) & 31u;
    w[12] = __float_as_int(rg / static_cast<float>(static_cast<int32_t>(1u << sh)));
  }
  __device__ __forceinline__ void children(const View& v, const SlotChildren& sc) {
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = key_value(v, sc.key(c));
    w[13] = static_cast<int32_t>(sc.node_mask);
    w[14] = 0;
  }
  __device__ __forceinline__ void store(OctNode* nd) const {
    int32_t* p = reinterpret_cast<int32_t*>(nd);
    const v4i a = {w[0], w[1], w[2], w[3]};
    const v4i b = {w[4], w[5], w[6], w[7]};
    const v4i c = {w[8], w[9], w[10], w[11]};
    const v3i d = {w[12], w[13], w[14]};
    __builtin_memcpy(p, &a, 16);
    __builtin_memcpy(p + 4, &b, 16);
    __builtin_memcpy(p + 8, &c, 16);
    __builtin_memcpy(p + 12, &d, 12);
  }
};

__device__ __forceinline__ void store_zeros(OctNode* nodes, int32_t from, int32_t to) {
  NodeOut z;
  z.zero();
  for (int32_t s = from; s < to; ++s) z.store(nodes + s);
}

__device__ __forceinline__ void build_root(const View& v, int32_t root_level) {
  const SegBind& b = v.b;
  SlotChildren sc;
  sc.init(v.sc_col);
  if (b.m >= 1) {
    // Node 0's own leaves link into the root slot.
    const Rec r0 = v.R(0);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (r0.leaf(side)) {
        const int32_t leaf = r0.child() + side;
        if (leaf >= 0 && leaf <= b.m)
          sc.put(child_of(v.CODE(leaf), r0.pn / 3 + 1), link_key(0, side));
      }
    }
    // The group grows through the true root's internal children, the nodes
    // whose parent entry is 0: the split of [0, n] (04_radix_tree.cpp:94-104
    // with d = +1, l = n).
    const int32_t n = b.m;
    const uint32_t c0 = v.CODE(0);
    const uint32_t dn = delta(c0, v.CODE(n));
    const int32_t lo = 31 - static_cast<int32_t>(__clz(n));
    const int32_t lg = lo + ((1u << lo) < static_cast<uint32_t>(n));
    int32_t s = 0;
    for (int32_t k = 1; k <= lg; ++k) {  // step ceil(n / 2^k), as a shift
      const int32_t st = s + static_cast<int32_t>((static_cast<uint32_t>(n) + (1u << k) - 1u) >> k);
      if (st <= n && delta(c0, v.CODE(st)) > dn) s = st;
    }
    int32_t mem[2] = {-1, -1};
    Rec mr[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int32_t y = s + q;
      const bool internal = q == 0 ? (s != 0) : (s + 1 != n);
      if (!internal || y <= 0 || y >= n) continue;
      const Rec ry = v.R(y);
      if (ry.e == 0) {
        mem[q] = y;
        mr[q] = ry;
      } else if (ry.e > 0) {
        const uint32_t w = child_of(v.CODE(y), ry.pn / 3 - ry.e + 1);
        sc.put(w, make_key(y, ry.e - 1));
        sc.node_mask |= 1u << w;
      }
    }
    // Members at depth 1 and below, down to depth 4.
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (mem[q] >= 0) walk_group(v, mem[q], mr[q], 1, 4, sc);
  }
  NodeOut nd;
  nd.children(v, sc);
  const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * root_level) & 31u;
  const uint32_t c = (v.CODE(0) >> shift) << shift;
  nd.w[8] = __float_as_int(decode_axis(c, 0, b.mn, b.rg));
  nd.w[9] = __float_as_int(decode_axis(c, 1, b.mn, b.rg));
  nd.w[10] = __float_as_int(decode_axis(c, 2, b.mn, b.rg));
  nd.w[11] = __float_as_int(1.0f);
  nd.w[12] = __float_as_int(b.rg);
  nd.store(b.nodes);
  // Slots before the first chain are never written by the reference.
  int32_t first = b.k;
  for (int32_t q = 1; q < b.m; ++q) {
    if (v.E(q) > 0) {
      first = v.P(q);
      break;
    }
  }
  store_zeros(b.nodes, 1, min(first, b.k));
}

// Child-index bits of step links into slot j (1..9) of a chain.
struct StepBits {
  uint64_t lo = 0;  // slots 0..7, 8 bits each
  uint32_t hi = 0;  // slots 8..9
  __device__ __forceinline__ void set(int32_t j, uint32_t w) {
    if (j < 8) lo |= static_cast<uint64_t>(1u << w) << (8 * j);
    else hi |= (1u << w) << (8 * (j - 8));
  }
  __device__ __forceinline__ uint32_t get(int32_t j) const {
    return j < 8 ? static_cast<uint32_t>(lo >> (8 * j)) & 0xffu
                 : (hi >> (8 * (j - 8))) & 0xffu;
  }
};

__device__ __forceinline__ void build_chain(const View& v, int32_t i, int32_t root_level,
                                            int32_t emax_win, int32_t reach_before) {
  const SegBind& b = v.b;
  const Rec ri = v.R(i);
  const int32_t nn = min(ri.e, kMaxEdge);
  const int32_t p_i = ri.p;
  const int32_t last = p_i + nn - 1;

  // Later chains: which slots stay ours, and where the next chain starts.
  uint32_t owned = (1u << nn) - 1u;
  int32_t next_start = b.k;
  bool found_next = false;
  auto fwd = [&](int32_t pq, int32_t eq) {  // one step; true ends the scan
    if (found_next && pq > last) return true;
    if (eq <= 0) return false;
    if (!found_next) {
      next_start = pq;
      found_next = true;
    }
    if (pq > last) return true;
    // slots max(pq, p_i) .. min(pq + eq - 1, last) pass to chain q: at most
    // kMaxEdge bits of `owned`
    const int32_t lo_s = max(pq, p_i);
    const int32_t cnt = min(pq + eq - 1, last) - lo_s + 1;
    if (cnt > 0) owned &= ~(((1u << cnt) - 1u) << (lo_s - p_i));
    return false;
  };
  {
    // Inside the window the scan reads one LDS record per step; past its
    // end it continues with the general lookups.
    int32_t q = i + 1;
    bool ended = false;
    for (const int32_t q_win = min(b.m, v.lo + v.n); q < q_win; ++q) {
      const v2i r = v.pe[q - v.lo];
      if (fwd(r.x, epn_e(r.y))) {
        ended = true;
        break;
      }
    }
    if (!ended)
      for (; q < b.m; ++q)
        if (fwd(v.P(q), v.E(q))) break;
  }

  // Step links into this chain's slots: its own, and those of earlier chains
  // that still cover them; their furthest end bounds the following gap.
  // Chains before the window end by reach_before; chains inside it by their
  // start plus the window's largest edge count.
  const uint32_t code_i = v.CODE(i);
  const int32_t lvl_i = ri.pn / 3;
  StepBits bits;
  SlotChildren sc;
  sc.init(v.sc_col);
  for (int32_t t = 0; t + 1 < nn; ++t) bits.set(t + 1, child_of(code_i, lvl_i - t));
  int32_t reach = p_i + nn;
  const bool before_matters = reach_before > p_i;
  // Links of chain q (first slot pq, eq steps) that cover slots of chain i.
  auto back_links = [&](int32_t q, int32_t pq, int32_t eq, uint32_t code_q, int32_t lvl_q) {
    const int32_t s_hi = min(pq + eq - 1, last);
    for (int32_t s = max(p_i, pq + 1); s <= s_hi; ++s) {
      const int32_t t = s - pq - 1;  // chain q's step t writes slot pq+t+1
      const uint32_t w = child_of(code_q, lvl_q - t);
      if (s == p_i) {
        sc.put(w, make_key(q, t));
        sc.node_mask |= 1u << w;
      } else {
        bits.set(s - p_i, w);
      }
    }
  };
  {
    int32_t q = i - 1;
    bool ended = false;
    for (const int32_t q_win = max(1, v.lo); q >= q_win; --q) {  // inside the window
      const v2i r = v.pe[q - v.lo];
      const int32_t pq = r.x;
      if (pq + emax_win <= p_i && !before_matters) {
        ended = true;
        break;
      }
      const int32_t eq = min(epn_e(r.y), kMaxEdge);
      if (eq <= 0) continue;
      reach = max(reach, pq + eq);
      if (pq + eq - 1 < p_i) continue;
      back_links(q, pq, eq, static_cast<uint32_t>(v.lcc[q - v.lo].y), epn_pn(r.y) / 3);
    }
    if (!ended) {
      for (; q >= 1; --q) {  // before the window
        const int32_t pq = v.P(q);
        if (pq + kMaxEdge <= p_i) break;
        const int32_t eq = min(v.E(q), kMaxEdge);
        if (eq <= 0) continue;
        reach = max(reach, pq + eq);
        if (pq + eq - 1 < p_i) continue;
        back_links(q, pq, eq, v.CODE(q), v.R(q).pn / 3);
      }
    }
  }

  // First slot: the earlier chains' step links, then the group's top links
  // and leaf links.
  walk_group(v, i, ri, 0, 3, sc);
  if (p_i > 0 && p_i < b.k) {
    NodeOut nd;
    nd.children(v, sc);
    nd.geometry(code_i, lvl_i, root_level, b.mn, b.rg);
    nd.store(b.nodes + p_i);
  }
  for (int32_t j = 1; j < nn; ++j) {
    if (!((owned >> j) & 1u)) continue;
    const int32_t s = p_i + j;
    if (s <= 0 || s >= b.k) continue;
    NodeOut nd;
    const uint32_t w = bits.get(j);
#pragma unroll
    for (int c = 0; c < 8; ++c) nd.w[c] = ((w >> c) & 1u) ? s - 1 : 0;
    nd.w[13] = static_cast<int32_t>(w);
    nd.w[14] = 0;
    nd.geometry(code_i, lvl_i - j, root_level, b.mn, b.rg);
    nd.store(b.nodes + s);
  }
  store_zeros(b.nodes, max(reach, 1), min(next_start, b.k));
}

// Eight waves per SIMD: left to itself the compiler gives the kernel 106
// scalar registers, which caps it at seven (800 per SIMD); asked for eight
// it keeps 78 and parks the rest in VGPR lanes. At 10M uniform points:
// 276 -> 254 us together with the 16-byte window records.
// BY_GROUP: order the tile's chains by the size of their group walk (see
// the compaction below); otherwise index order. Which is faster depends on
// the cloud: at 10M uniform points index order (253 vs 259 us), at 2^24
// clustered points group order (318 vs 341 us); the autotuner chooses.
template <bool BY_GROUP>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void oct_build_kernel(
    OctArgs a, SegView sv, const int32_t* __restrict__ tile_start) {
  __shared__ v2i s_pe[kWinN];
  __shared__ v2i s_lcc[kWinN];
  static_assert(kNodeTile <= 256, "work items are stored as bytes");
  __shared__ uint8_t s_list[kNodeTile];  // work items as i - i0
  __shared__ uint32_t s_sc[8 * kThreads];
  __shared__ int32_t s_count;
  __shared__ int32_t s_count_b;  // group order: work items placed from the list's end
  __shared__ int32_t s_emax;

  const TileRef r = find_tile(tile_start, sv.n_seg, xcd_tile(blockIdx.x, gridDim.x));
  // The segment's scalars are loaded together before any is tested: one
  // round trip instead of one per value.
  // (The launcher guarantees these pointers; the build writes them all.)
  const uint32_t fits = a.ctrl[kCtrlOctOk];
  const int32_t m_seg = a.n_unique[r.seg] - 1;
  const int32_t node_base = a.oct_start[r.seg];
  const int32_t k_seg = a.kalloc[r.seg];
  const int32_t sb = seg_begin(sv, r.seg);
  const float seg_mn = seg_min(sv, r.seg);
  const float seg_rg = seg_range(sv, r.seg);
  const int32_t i0 = r.local * kNodeTile;
  // One test after all the loads (an early return per value would hold the
  // loads after it back): the nodes fit, the tile has nodes (a single-key
  // segment still has its root), the segment has slots and a valid base.
  // After a lookback time-out the inputs are not valid.
  const uint32_t stop = a.ctrl[kCtrlTimeout];
  if (fits == 0 || stop != 0 || (i0 >= m_seg && !(i0 == 0 && m_seg == 0)) || k_seg < 1 ||
      node_base < 0)
    return;
  SegBind b;
  b.m = m_seg;
  b.nodes = a.nodes + node_base;
  b.k = k_seg;
  b.codes = a.codes + sb;
  b.pn = a.prefix_n + sb;
  b.hl = a.has_leaf_left + sb;
  b.hr = a.has_leaf_right + sb;
  b.lc = a.left_child + sb;
  b.e = a.edge + sb;
  b.p = a.offsets + sb;
  b.mn = seg_mn;
  b.rg = seg_rg;

  const int32_t lo = max(0, i0 - kHalo);
  const int32_t wn = max(0, min(b.m, i0 + kNodeTile + kHalo) - lo);
  const int32_t wnc = max(0, min(b.m + 1, i0 + kNodeTile + kHalo) - lo);
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_count = 0;
    if (BY_GROUP) s_count_b = 0;
    s_emax = 0;
  }
  __syncthreads();
  // All of the window's loads are issued before any is used: one round
  // trip to memory instead of one per row of the window.
  constexpr int kRows = (kWinN + kThreads - 1) / kThreads;
  int32_t emax = 0;
  uint32_t r_code[kRows];
  int32_t r_p[kRows], r_e[kRows], r_lc[kRows];
  uint8_t r_hl[kRows], r_hr[kRows], r_pn[kRows];
#pragma unroll
  for (int k = 0; k < kRows; ++k) {
    const int32_t q = tid + k * kThreads;
    const int32_t g = lo + q;
    if (q < wnc) r_code[k] = b.codes[g];
    if (q < wn) {
      r_p[k] = b.p[g];
      r_e[k] = b.e[g];
      r_lc[k] = b.lc[g];
      r_hl[k] = b.hl[g];
      r_hr[k] = b.hr[g];
      r_pn[k] = b.pn[g];
    }
  }
#pragma unroll
  for (int k = 0; k < kRows; ++k) {
    const int32_t q = tid + k * kThreads;
    if (q < wn) {
      const uint32_t l = static_cast<uint32_t>(r_lc[k]) | (r_hl[k] ? kLeafL : 0u) |
                         (r_hr[k] ? kLeafR : 0u);
      s_pe[q] = v2i{r_p[k], pack_epn(r_e[k], r_pn[k])};
      s_lcc[q] = v2i{static_cast<int32_t>(l), static_cast<int32_t>(r_code[k])};
      emax = max(emax, r_e[k]);
    } else if (q < wnc) {
      s_lcc[q] = v2i{0, static_cast<int32_t>(r_code[k])};
    }
  }
  emax = wave_incl_max(emax);
  if (lane_id() == 0) atomicMax(&s_emax, emax);
  __syncthreads();
  const View v{b, lo, wn, wnc, as_lds<const v2i>(s_pe), as_lds<const v2i>(s_lcc),
               as_lds<uint32_t>(s_sc + tid)};
  // Compact the block's work: the root and the non-empty chains. In group
  // order, chains none of whose own children is an e == 0 internal node (a
  // group walk of one member) fill the list from the front and the others
  // (a child outside the window counts as such, as does the root) from the
  // back, so waves of the first kind walk one member and the rest walk
  // their like: at 10M uniform points the slowest lane of a wave walks 4.5
  // members on average instead of 6.9 in index order.
  constexpr int kH = kNodeTile / kThreads;
#pragma unroll
  for (int h = 0; h < kH; ++h) {
    const int32_t i = i0 + h * kThreads + tid;
    const bool work = (i == 0) || (i < b.m && v.E(i) > 0);
    bool back = false;
    if (BY_GROUP && work) {
      if (i == 0) {
        back = true;
      } else {
        const Rec ri = v.R(i);
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int32_t ch = ri.child() + side;
          if (ri.leaf(side) || ch <= 0 || ch >= b.m) continue;
          if (static_cast<uint32_t>(ch - v.lo) >= static_cast<uint32_t>(v.n) || v.E(ch) == 0)
            back = true;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < (BY_GROUP ? 2 : 1); ++k) {
      const bool sel = work && (k == 1) == back;
      const uint64_t bal = __ballot(sel);
      uint32_t base = 0;
      if (lane_id() == 0 && bal)
        base = atomicAdd(reinterpret_cast<uint32_t*>(k == 0 ? &s_count : &s_count_b),
                         static_cast<uint32_t>(__popcll(bal)));
      base = __shfl(base, 0, 64);
      const uint32_t pos = base + __popcll(bal & lanemask_lt());
      if (sel) s_list[k == 0 ? pos : kNodeTile - 1 - pos] = static_cast<uint8_t>(i - i0);
    }
  }
  __syncthreads();
  const int32_t root_level = b.m >= 1 ? v.R(0).pn / 3 : 0;
  const int32_t emax_win = min(s_emax, kMaxEdge);
  // Chains before the window start no later than P[lo] and end within
  // kMaxEdge slots of their start.
  const int32_t reach_before = (lo > 0 && wn > 0) ? v.P(lo) + kMaxEdge : 0;
  const int32_t cnt_a = s_count;
  const int32_t cnt = BY_GROUP ? cnt_a + s_count_b : cnt_a;
  for (int q = tid; q < cnt; q += kThreads) {
    // (group order: the back items sit at the end of the list)
    const int32_t slot = (!BY_GROUP || q < cnt_a) ? q : kNodeTile - cnt + q;
    const int32_t node = i0 + static_cast<int32_t>(s_list[slot]);
    if (node == 0) build_root(v, root_level);
    else build_chain(v, node, root_level, emax_win, reach_before);
  }
}

// Root node of apps/cpu_demo/main.cpp:125-132 alone (the per-stage entry
// point that stands in for that host code).
__global__ void oct_root_kernel(OctNode* nodes, const uint32_t* codes,
                                const uint8_t* prefix_n, int32_t n_brt,
                                float mn, float rg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t root_level = n_brt >= 1 ? prefix_n[0] / 3 : 0;
  const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * root_level) & 31u;
  const uint32_t c = (codes[0] >> shift) << shift;
  nodes->corner[0] = decode_axis(c, 0, mn, rg);
  nodes->corner[1] = decode_axis(c, 1, mn, rg);
  nodes->corner[2] = decode_axis(c, 2, mn, rg);
  nodes->corner[3] = 1.0f;
  nodes->cell_size = rg;
}

}  // namespace

void launch_oct_root(OctNode* nodes, const uint32_t* codes,
                     const uint8_t* prefix_n, int32_t n_brt, float mn,
                     float rg, hipStream_t s) {
  oct_root_kernel<<<1, 64, 0, s>>>(nodes, codes, prefix_n, n_brt, mn, rg);
}

bool launch_oct_build(const OctArgs& a, const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles, bool by_group,
                      hipStream_t s) {
  // The group walk reads the children arrays; the per-segment counts, node
  // bases and the fit flag come from the build.
  if (!a.has_leaf_left || !a.has_leaf_right || !a.left_child || !a.codes ||
      !a.prefix_n || !a.edge || !a.offsets || !a.nodes || !a.ctrl || !a.n_unique ||
      !a.oct_start || !a.kalloc)
    return false;
  if (total_tiles <= 0) return true;
  if (by_group)
    oct_build_kernel<true><<<total_tiles, kThreads, 0, s>>>(a, sv, tile_start);
  else
    oct_build_kernel<false><<<total_tiles, kThreads, 0, s>>>(a, sv, tile_start);
  return true;
}

}  // namespace ocp

/*
 * CPU oracle for the seven-stage octree pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * of bench.py. The product path never links or calls this file.
 *
 * Canonical semantics: every function below is the reference CPU function
 * (file:line cited per function) run single-threaded, in ascending index
 * order, on zero-initialised outputs, with these resolutions of undefined
 * behaviour in the reference:
 *   - clz(0) := 32, so delta(a, a) = 31 and log2_ceil(0) := 0
 *     (04_radix_tree.cpp:26,32; reached only when U == 2);
 *   - writes to parent[] at an index >= n are dropped (04_radix_tree.cpp:117);
 *   - parent[0] is 0 (the reference relies on zeroed memory);
 *   - shift counts are taken mod 32, as x86 shr/shl and AMD v_lshrrev do
 *     (oct.h:106,122 reach 30-3*11 = -3 for leaf_level 11; oct.h:63 for a
 *     level below the root level);
 *   - the leaf-link ancestor walk (oct.h:111-113,126-128) stops at the root,
 *     where the reference would spin forever, and after 64 hops;
 *   - slot indices outside [0, n_alloc) are not written.
 * The OpenMP variants keep the reference's loop structure and are used only
 * for timing: with more than one thread the octree stage races exactly as the
 * reference does (oct.h:12-13).
 */
#include "octree_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int kMortonBits = 30;  // include/shared/morton.h:7

inline uint32_t clz32(uint32_t x) { return x ? __builtin_clz(x) : 32u; }

inline uint32_t split_by_3(uint32_t a) {
  // Magic-bit spread of a 10-bit field to every third bit, morton.h:22-29.
  uint32_t x = a & 0x3ffu;
  x = (x | x << 16) & 0x30000ffu;
  x = (x | x << 8) & 0x0300f00fu;
  x = (x | x << 4) & 0x30c30c3u;
  x = (x | x << 2) & 0x9249249u;
  return x;
}

inline uint32_t third_bits(uint32_t m) {
  // Inverse of split_by_3, morton.h:52-59.
  uint32_t x = m & 0x9249249u;
  x = (x ^ (x >> 2)) & 0x30c30c3u;
  x = (x ^ (x >> 4)) & 0x0300f00fu;
  x = (x ^ (x >> 8)) & 0x30000ffu;
  x = (x ^ (x >> 16)) & 0x3ffu;
  return x;
}

inline uint32_t quantize(float a, float min_coord, float range) {
  // morton.h:42-44: float sub, float div, times int 1024, truncating cast.
  const float v = (a - min_coord) / range * 1024;
  return static_cast<uint32_t>(static_cast<int64_t>(v));
}

inline uint32_t encode(float x, float y, float z, float min_coord,
                       float range) {
  return split_by_3(quantize(x, min_coord, range)) |
         (split_by_3(quantize(y, min_coord, range)) << 1) |
         (split_by_3(quantize(z, min_coord, range)) << 2);
}

inline void decode(uint32_t code, float min_coord, float range, float* out) {
  // morton.h:67-87; the products and sums are separately rounded.
  const float bit_scale = 1024.0f;
  volatile float t;
  for (int a = 0; a < 3; ++a) {
    t = static_cast<float>(third_bits(code >> a)) / bit_scale;
    t = t * range;
    out[a] = t + min_coord;
  }
  out[3] = 1.0f;
}

inline uint32_t delta(uint32_t a, uint32_t b) {
  // 04_radix_tree.cpp:21-27, returned through uint8_t.
  return static_cast<uint8_t>(clz32(a ^ b) - 1u);
}

inline int32_t log2_ceil(uint32_t x) {
  // 04_radix_tree.cpp:29-37.
  if (x == 0) return 0;
  const int32_t n_lower_bits = 31 - static_cast<int32_t>(clz32(x));
  return n_lower_bits + ((1u << n_lower_bits) < x);
}

inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

void radix_tree_node(int32_t i, int32_t n, const uint32_t* codes,
                     uint8_t* prefix_n, uint8_t* has_leaf_left,
                     uint8_t* has_leaf_right, int32_t* left_child,
                     int32_t* parent) {
  // Karras 2012 as written in 04_radix_tree.cpp:47-120.
  const uint32_t code_i = codes[i];
  int32_t d;
  if (i == 0) {
    d = 1;
  } else {
    const int32_t right = static_cast<int32_t>(delta(code_i, codes[i + 1]));
    const int32_t left = static_cast<int32_t>(delta(code_i, codes[i - 1]));
    const int32_t diff = right - left;
    d = (diff > 0) - (diff < 0);
  }
  int32_t l = 0;
  if (i == 0) {
    l = n - 1;
  } else {
    const uint32_t delta_min = delta(code_i, codes[i - d]);
    int64_t l_max = 2;
    while (i + l_max * d >= 0 && i + l_max * d <= n &&
           delta(code_i, codes[i + l_max * d]) > delta_min &&
           l_max < (int64_t(1) << 31)) {
      l_max *= 2;
    }
    const int32_t l_cutoff = (d == -1) ? i : n - i;
    for (int64_t t = l_max / 2, divisor = 2; t >= 1;
         divisor *= 2, t = l_max / divisor) {
      if (l + t <= l_cutoff &&
          delta(code_i, codes[i + (l + t) * d]) > delta_min) {
        l += static_cast<int32_t>(t);
      }
    }
  }
  const int32_t j = i + l * d;
  const uint32_t delta_node = delta(codes[i], codes[j]);
  prefix_n[i] = static_cast<uint8_t>(delta_node);
  int32_t s = 0;
  const int64_t max_divisor = int64_t(1) << log2_ceil(static_cast<uint32_t>(l));
  const int32_t s_cutoff = (d == -1) ? i - 1 : n - i - 1;
  int64_t divisor = 2;
  for (uint32_t t = ceil_div(static_cast<uint32_t>(l), 2); divisor <= max_divisor;
       divisor <<= 1, t = ceil_div(static_cast<uint32_t>(l),
                                   static_cast<uint32_t>(divisor))) {
    if (s + static_cast<int32_t>(t) <= s_cutoff &&
        delta(code_i, codes[i + (s + static_cast<int32_t>(t)) * d]) >
            delta_node) {
      s += static_cast<int32_t>(t);
    }
  }
  const int32_t gamma = i + s * d + std::min(d, 0);
  left_child[i] = gamma;
  has_leaf_left[i] = (std::min(i, j) == gamma);
  has_leaf_right[i] = (std::max(i, j) == gamma + 1);
  if (!has_leaf_left[i] && gamma >= 0 && gamma < n) parent[gamma] = i;
  if (!has_leaf_right[i] && gamma + 1 >= 0 && gamma + 1 < n)
    parent[gamma + 1] = i;
}

inline void set_child(orc_oct_node* nodes, int32_t n_alloc, int32_t slot,
                      uint32_t which, int32_t value) {
  // oct.h:9-14.
  if (slot < 0 || slot >= n_alloc) return;
  nodes[slot].children[which] = value;
  nodes[slot].child_node_mask |= 1 << which;
}

inline void set_leaf(orc_oct_node* nodes, int32_t n_alloc, int32_t slot,
                     uint32_t which, int32_t value) {
  // oct.h:16-21: note the AND-NOT on the leaf mask.
  if (slot < 0 || slot >= n_alloc) return;
  nodes[slot].children[which] = value;
  nodes[slot].child_leaf_mask &= ~(1 << which);
}

inline void node_geometry(orc_oct_node* nodes, int32_t n_alloc, int32_t slot,
                          uint32_t code, int32_t level, int32_t root_level,
                          float min_coord, float range) {
  // oct.h:44-57 and oct.h:76-88: corner of the level-`level` prefix and
  // range / 2^(level - root_level).
  if (slot < 0 || slot >= n_alloc) return;
  const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * level) & 31u;
  const uint32_t prefix = code >> shift;
  decode(prefix << shift, min_coord, range, nodes[slot].corner);
  const uint32_t sh = static_cast<uint32_t>(level - root_level) & 31u;
  const int32_t denom = static_cast<int32_t>(1u << sh);
  volatile float cs = range / static_cast<float>(denom);
  nodes[slot].cell_size = cs;
}

void process_oct_node(int32_t i, orc_oct_node* nodes, const int32_t* offsets,
                      const int32_t* counts, const uint32_t* codes,
                      const uint8_t* prefix_n, const int32_t* parent,
                      float min_coord, float range, int32_t n_alloc) {
  // oct.h:24-90.
  int32_t oct_idx = offsets[i];
  const int32_t n_new = counts[i];
  const int32_t root_level = prefix_n[0] / 3;
  for (int32_t j = 0; j < n_new - 1; ++j) {
    const int32_t level = prefix_n[i] / 3 - j;
    const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * level) & 31u;
    const uint32_t which = (codes[i] >> shift) & 7u;
    const int32_t up = oct_idx + 1;
    set_child(nodes, n_alloc, up, which, oct_idx);
    node_geometry(nodes, n_alloc, oct_idx, codes[i], level, root_level,
                  min_coord, range);
    oct_idx = up;
  }
  if (n_new > 0) {
    int32_t rt_parent = parent[i];
    int32_t counter = 0;
    while (counts[rt_parent] == 0) {
      rt_parent = parent[rt_parent];
      ++counter;
      if (counter > 30) break;
    }
    const int32_t oct_parent = offsets[rt_parent];
    const int32_t top_level = prefix_n[i] / 3 - n_new + 1;
    const uint32_t shift =
        static_cast<uint32_t>(kMortonBits - 3 * top_level) & 31u;
    const uint32_t which = (codes[i] >> shift) & 7u;
    set_child(nodes, n_alloc, oct_parent, which, oct_idx);
    node_geometry(nodes, n_alloc, oct_idx, codes[i], top_level, root_level,
                  min_coord, range);
  }
}

inline int32_t link_walk(int32_t node, const int32_t* counts,
                         const int32_t* parent) {
  // oct.h:111-113: first ancestor-or-self with a nonzero edge count; stops at
  // the root (slot 0) and after 64 hops.
  int32_t hops = 0;
  while (counts[node] == 0 && node != 0 && hops < 64) {
    node = parent[node];
    ++hops;
  }
  return node;
}

void process_link_leaf(int32_t i, orc_oct_node* nodes, const int32_t* offsets,
                       const int32_t* counts, const uint32_t* codes,
                       const uint8_t* has_leaf_left,
                       const uint8_t* has_leaf_right, const uint8_t* prefix_n,
                       const int32_t* parent, const int32_t* left_child,
                       int32_t n_alloc) {
  // oct.h:92-134.
  for (int side = 0; side < 2; ++side) {
    const bool has = side == 0 ? has_leaf_left[i] : has_leaf_right[i];
    if (!has) continue;
    const int32_t leaf_idx = left_child[i] + side;
    const int32_t leaf_level = prefix_n[i] / 3 + 1;
    const uint32_t shift =
        static_cast<uint32_t>(kMortonBits - 3 * leaf_level) & 31u;
    const uint32_t which = (codes[leaf_idx] >> shift) & 7u;
    const int32_t rt_node = link_walk(i, counts, parent);
    set_leaf(nodes, n_alloc, offsets[rt_node], which, leaf_idx);
  }
}

int threads_or_max(int nthreads) {
  return nthreads > 0 ? nthreads : omp_get_max_threads();
}

// The OpenMP LSD radix sort of kernels/02_sort.cpp:13-58: per-thread digit
// counts, merged under a critical section, one serial scan, thread-ordered
// bucket offsets, then a stable per-thread scatter.
void omp_radix_sort(uint32_t* data, int64_t n, int nthreads) {
  uint32_t* buffer =
      static_cast<uint32_t*>(std::malloc(static_cast<size_t>(n) * sizeof(uint32_t)));
  uint32_t* src = data;
  for (int shift = 0; shift < 32; shift += 8) {
    int64_t bucket[256] = {0};
    std::vector<int64_t> locals(static_cast<size_t>(nthreads) * 256, 0);
#pragma omp parallel num_threads(nthreads)
    {
      const int tid = omp_get_thread_num();
      const int nt = omp_get_num_threads();
      int64_t* local = &locals[static_cast<size_t>(tid) * 256];
#pragma omp for schedule(static) nowait
      for (int64_t i = 0; i < n; i++) local[(src[i] >> shift) & 255]++;
#pragma omp critical
      for (int b = 0; b < 256; b++) bucket[b] += local[b];
#pragma omp barrier
#pragma omp single
      for (int b = 1; b < 256; b++) bucket[b] += bucket[b - 1];
      for (int cur = nt - 1; cur >= 0; cur--) {
        if (cur == tid) {
          for (int b = 0; b < 256; b++) {
            bucket[b] -= local[b];
            local[b] = bucket[b];
          }
        }
#pragma omp barrier
      }
#pragma omp for schedule(static)
      for (int64_t i = 0; i < n; i++)
        buffer[local[(src[i] >> shift) & 255]++] = src[i];
    }
    std::swap(src, buffer);
  }
  // Four passes: the result is back in `data` and `buffer` is the scratch.
  std::free(buffer);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

uint32_t orc_split_by_3(uint32_t a) { return split_by_3(a); }
uint32_t orc_encode(float x, float y, float z, float min_coord, float range) {
  return encode(x, y, z, min_coord, range);
}
uint32_t orc_third_bits(uint32_t m) { return third_bits(m); }
void orc_decode(uint32_t code, float min_coord, float range, float out[4]) {
  decode(code, min_coord, range, out);
}
uint32_t orc_delta(uint32_t a, uint32_t b) { return delta(a, b); }
int32_t orc_log2_ceil(uint32_t x) { return log2_ceil(x); }
int orc_max_threads(void) { return omp_get_num_procs(); }

void orc_morton(const float* pts, uint32_t* keys, int64_t n, float min_coord,
                float range, int nthreads) {
  // kernels/01_morton.cpp:5-14.
  const int nt = threads_or_max(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const float* p = pts + 4 * i;
    keys[i] = encode(p[0], p[1], p[2], min_coord, range);
  }
}

void orc_morton3(const float* pts, uint32_t* keys, int64_t n, float min_coord,
                 float range) {
  for (int64_t i = 0; i < n; i++) {
    const float* p = pts + 3 * i;
    keys[i] = encode(p[0], p[1], p[2], min_coord, range);
  }
}

void orc_radix_sort(uint32_t* keys, int64_t n, int nthreads) {
  // kernels/02_sort.cpp:110-112.
  if (n <= 1) return;
  omp_radix_sort(keys, n, threads_or_max(nthreads));
}

int32_t orc_count_unique(uint32_t* keys, int32_t n) {
  // kernels/03_unique.cpp:2-16.
  if (n <= 0) return 0;
  int32_t j = 0;
  for (int32_t i = 1; i < n; ++i) {
    if (keys[i] != keys[j]) {
      ++j;
      keys[j] = keys[i];
    }
  }
  return j + 1;
}

void orc_build_radix_tree(int32_t n, const uint32_t* codes, uint8_t* prefix_n,
                          uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                          int32_t* left_child, int32_t* parent, int nthreads) {
  // kernels/04_radix_tree.cpp:39-121, n BRT nodes over n+1 codes.
  if (n <= 0) return;
  std::memset(prefix_n, 0, n);
  std::memset(has_leaf_left, 0, n);
  std::memset(has_leaf_right, 0, n);
  std::memset(left_child, 0, sizeof(int32_t) * n);
  std::memset(parent, 0, sizeof(int32_t) * n);
  if (nthreads == 1) {
    for (int32_t i = 0; i < n; i++)
      radix_tree_node(i, n, codes, prefix_n, has_leaf_left, has_leaf_right,
                      left_child, parent);
  } else {
#pragma omp parallel for num_threads(threads_or_max(nthreads)) schedule(static)
    for (int32_t i = 0; i < n; i++)
      radix_tree_node(i, n, codes, prefix_n, has_leaf_left, has_leaf_right,
                      left_child, parent);
  }
}

void orc_edge_count(const uint8_t* prefix_n, const int32_t* parent,
                    int32_t* edge_count, int32_t n_brt, int nthreads) {
  // kernels/05_edge_count.cpp:3-16.
  if (n_brt <= 0) return;
  edge_count[0] = 0;
#pragma omp parallel for num_threads(threads_or_max(nthreads)) schedule(static)
  for (int32_t i = 1; i < n_brt; ++i) {
    const int32_t my_depth = prefix_n[i] / 3;
    const int32_t parent_depth = prefix_n[parent[i]] / 3;
    edge_count[i] = my_depth - parent_depth;
  }
}

void orc_partial_sum(const int32_t* in, int32_t n, int32_t* out) {
  // kernels/06_partial_sum.cpp:3-15 (std::partial_sum semantics).
  if (n <= 0) return;
  int32_t sum = in[0];
  out[0] = sum;
  for (int32_t i = 1; i < n; ++i) {
    sum += in[i];
    out[i] = sum;
  }
}

int32_t orc_oct_alloc_count(const int32_t* offsets, const int32_t* counts,
                            int32_t n_brt, int32_t n_unique) {
  if (n_unique <= 0) return 0;
  int32_t k = 1;
  for (int32_t i = 0; i < n_brt; ++i) k = std::max(k, offsets[i] + counts[i]);
  return k;
}

void orc_root_init(orc_oct_node* nodes, const uint32_t* codes,
                   const uint8_t* prefix_n, int32_t n_brt, float min_coord,
                   float range) {
  // apps/cpu_demo/main.cpp:125-132 (benchmarks/fixture.hpp:73-80).
  const int32_t root_level = n_brt > 0 ? prefix_n[0] / 3 : 0;
  const uint32_t shift = static_cast<uint32_t>(kMortonBits - 3 * root_level) & 31u;
  const uint32_t root_prefix = codes[0] >> shift;
  decode(root_prefix << shift, min_coord, range, nodes[0].corner);
  nodes[0].cell_size = range;
}

void orc_make_oct_nodes(orc_oct_node* nodes, const int32_t* offsets,
                        const int32_t* counts, const uint32_t* codes,
                        const uint8_t* prefix_n, const int32_t* parent,
                        float min_coord, float range, int32_t n_brt,
                        int32_t n_alloc, int nthreads) {
  // kernels/07_octree.cpp:5-31 (i from 1, N = number of BRT nodes).
  if (nthreads == 1) {
    for (int32_t i = 1; i < n_brt; ++i)
      process_oct_node(i, nodes, offsets, counts, codes, prefix_n, parent,
                       min_coord, range, n_alloc);
  } else {
#pragma omp parallel for num_threads(threads_or_max(nthreads)) schedule(static)
    for (int32_t i = 1; i < n_brt; ++i)
      process_oct_node(i, nodes, offsets, counts, codes, prefix_n, parent,
                       min_coord, range, n_alloc);
  }
}

void orc_link_leaf_nodes(orc_oct_node* nodes, const int32_t* offsets,
                         const int32_t* counts, const uint32_t* codes,
                         const uint8_t* has_leaf_left,
                         const uint8_t* has_leaf_right,
                         const uint8_t* prefix_n, const int32_t* parent,
                         const int32_t* left_child, int32_t n_brt,
                         int32_t n_alloc, int nthreads) {
  // kernels/07_octree.cpp:33-57.
  if (nthreads == 1) {
    for (int32_t i = 0; i < n_brt; ++i)
      process_link_leaf(i, nodes, offsets, counts, codes, has_leaf_left,
                        has_leaf_right, prefix_n, parent, left_child, n_alloc);
  } else {
#pragma omp parallel for num_threads(threads_or_max(nthreads)) schedule(static)
    for (int32_t i = 0; i < n_brt; ++i)
      process_link_leaf(i, nodes, offsets, counts, codes, has_leaf_left,
                        has_leaf_right, prefix_n, parent, left_child, n_alloc);
  }
}

int32_t orc_time_pipeline(const float* pts, int64_t n, float min_coord,
                          float range, int nthreads, int reps,
                          double* stage_ms, int32_t* out_k) {
  // The chain of apps/cpu_demo/main.cpp:63-155 with the canonical call
  // convention (BRT over U-1 nodes, edge count and scan over U-1 entries).
  const int nt = threads_or_max(nthreads);
  if (reps < 1) reps = 1;
  std::vector<uint32_t> keys(static_cast<size_t>(n));
  std::vector<std::vector<double>> t(8);
  int32_t u = 0, k = 0;
  for (int r = 0; r < reps + 1; ++r) {
    double ts[8];
    const double t0 = now_ms();
    orc_morton(pts, keys.data(), n, min_coord, range, nt);
    const double t1 = now_ms();
    orc_radix_sort(keys.data(), n, nt);
    const double t2 = now_ms();
    u = orc_count_unique(keys.data(), static_cast<int32_t>(n));
    const double t3 = now_ms();
    const int32_t m = std::max(u - 1, 0);
    std::vector<uint8_t> pn(m), hl(m), hr(m);
    std::vector<int32_t> lc(m), par(m), e(m), p(m);
    const double t3b = now_ms();
    orc_build_radix_tree(m, keys.data(), pn.data(), hl.data(), hr.data(),
                         lc.data(), par.data(), nt);
    const double t4 = now_ms();
    orc_edge_count(pn.data(), par.data(), e.data(), m, nt);
    const double t5 = now_ms();
    orc_partial_sum(e.data(), m, p.data());
    if (m > 0) p[0] = 0;
    const double t6 = now_ms();
    k = orc_oct_alloc_count(p.data(), e.data(), m, u);
    std::vector<orc_oct_node> nodes(static_cast<size_t>(std::max(k, 1)));
    const double t6b = now_ms();
    if (u > 0) orc_root_init(nodes.data(), keys.data(), pn.data(), m, min_coord, range);
    orc_make_oct_nodes(nodes.data(), p.data(), e.data(), keys.data(), pn.data(),
                       par.data(), min_coord, range, m, k, nt);
    orc_link_leaf_nodes(nodes.data(), p.data(), e.data(), keys.data(),
                        hl.data(), hr.data(), pn.data(), par.data(), lc.data(),
                        m, k, nt);
    const double t7 = now_ms();
    ts[0] = t1 - t0;
    ts[1] = t2 - t1;
    ts[2] = t3 - t2;
    ts[3] = t4 - t3b;
    ts[4] = t5 - t4;
    ts[5] = t6 - t5;
    ts[6] = t7 - t6b;
    ts[7] = ts[0] + ts[1] + ts[2] + ts[3] + ts[4] + ts[5] + ts[6];
    if (r > 0)  // the first run is a warm-up
      for (int s = 0; s < 8; ++s) t[s].push_back(ts[s]);
  }
  for (int s = 0; s < 8; ++s) {
    std::sort(t[s].begin(), t[s].end());
    stage_ms[s] = t[s][t[s].size() / 2];
  }
  if (out_k) *out_k = k;
  return u;
}

}  // extern "C"

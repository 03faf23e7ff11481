// octree_fixture.hpp - the GpuFixture of the reference's GPU benchmarks
// (benchmarks/cuda/fixture.cuh:24-91) over the C ABI of octree_pipeline.h,
// header-only C++17 with no HIP or google-benchmark dependency.
//
// SetUp() prepares every stage input except the octree, as the reference's
// fixture does: kN float4 points (glm::vec4 layout, w = 1) from std::mt19937
// seeded with kRandomSeed through std::uniform_real_distribution(kMin, kMax),
// drawn x, y, z in order; Morton codes; the sorted keys uniqued in place
// (u_sort, num_unique); the radix tree over num_unique - 1 nodes (tree);
// edge counts; their inclusive scan with u_prefix_sum[0] = 0. The members
// have the reference's names. Differences, each deliberate:
//  - every array is device memory (the reference uses managed memory): read
//    it on the host with ocp_memcpy_d2h;
//  - every stage runs on the device (the reference sorts, uniques and scans
//    on the host);
//  - num_oct_nodes is set: max(1, max_i u_prefix_sum[i] + u_edge_count[i]),
//    the slot count the octree entries take as n_alloc (the reference
//    leaves it unset);
//  - errors throw std::runtime_error with ocp_last_error's text instead of
//    exiting.
#pragma once

#include <algorithm>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "octree_pipeline.h"

namespace ocp {

// benchmarks/cuda/config.cuh:4-8.
constexpr int kN = 10'000'000;
constexpr float kMin = 0.0f;
constexpr float kMax = 1024.0f;
constexpr float kRange = kMax - kMin;
constexpr unsigned kRandomSeed = 114514;

// include/shared/types.h:14-21; the booleans are bytes, as in the ABI.
struct RadixTreeData {
  int n_nodes = 0;
  uint8_t* prefixN = nullptr;
  uint8_t* hasLeafLeft = nullptr;
  uint8_t* hasLeafRight = nullptr;
  int* leftChild = nullptr;
  int* parent = nullptr;
};

class GpuFixture {
 public:
  explicit GpuFixture(int n = kN, int device = 0) : n_(n), device_(device) {}
  GpuFixture(const GpuFixture&) = delete;
  GpuFixture& operator=(const GpuFixture&) = delete;
  ~GpuFixture() {
    TearDown();
    if (ctx) ocp_destroy(ctx);
  }

  void SetUp() {
    TearDown();
    if (!ctx) check(ocp_create(&ctx, device_, nullptr), "ocp_create");
    std::vector<float> host(static_cast<size_t>(n_) * 4);
    std::mt19937 gen(kRandomSeed);
    std::uniform_real_distribution<float> dis(kMin, kMax);
    for (int i = 0; i < n_; ++i) {
      host[4 * size_t(i) + 0] = dis(gen);
      host[4 * size_t(i) + 1] = dis(gen);
      host[4 * size_t(i) + 2] = dis(gen);
      host[4 * size_t(i) + 3] = 1.0f;
    }
    u_points = alloc<float>(size_t(n_) * 4);
    check(ocp_memcpy_h2d(ctx, u_points, host.data(), host.size() * sizeof(float)), "upload");

    u_sort = alloc<unsigned int>(n_);
    unsigned int* alt = alloc<unsigned int>(n_);
    int32_t* d_count = alloc<int32_t>(1);
    check(ocp_k_ComputeMortonCode(ctx, u_points, u_sort, size_t(n_), kMin, kRange), "morton");
    check(ocp_k_SortKeys(ctx, u_sort, alt, uint32_t(n_)), "sort");
    check(ocp_k_CountUnique(ctx, u_sort, d_count, n_), "unique");
    int32_t u = 0;
    check(ocp_memcpy_d2h(ctx, &u, d_count, sizeof(u)), "unique count");
    num_unique = u;

    tree.n_nodes = std::max(num_unique - 1, 0);
    const size_t m = size_t(tree.n_nodes);
    tree.prefixN = alloc<uint8_t>(m);
    tree.hasLeafLeft = alloc<uint8_t>(m);
    tree.hasLeafRight = alloc<uint8_t>(m);
    tree.leftChild = alloc<int>(m);
    tree.parent = alloc<int>(m);
    check(ocp_k_BuildRadixTree(ctx, tree.n_nodes, u_sort, tree.prefixN, tree.hasLeafLeft,
                               tree.hasLeafRight, tree.leftChild, tree.parent),
          "radix tree");

    u_edge_count = alloc<int>(m);
    u_prefix_sum = alloc<int>(m + 1);
    check(ocp_k_EdgeCount(ctx, tree.prefixN, tree.parent, u_edge_count, tree.n_nodes),
          "edge count");
    check(ocp_k_PartialSum(ctx, u_edge_count, 0, tree.n_nodes, u_prefix_sum), "partial sum");
    check(ocp_memset(ctx, u_prefix_sum, 0, sizeof(int)), "prefix_sum[0]");

    std::vector<int> e(m), p(m);
    if (m) {
      check(ocp_memcpy_d2h(ctx, e.data(), u_edge_count, m * sizeof(int)), "edge read");
      check(ocp_memcpy_d2h(ctx, p.data(), u_prefix_sum, m * sizeof(int)), "scan read");
    }
    int slots = 1;
    for (size_t i = 0; i < m; ++i) slots = std::max(slots, p[i] + e[i]);
    num_oct_nodes = slots;
    check(ocp_synchronize(ctx), "set-up");
  }

  void TearDown() {
    for (void* ptr : owned_) ocp_free(ctx, ptr);
    owned_.clear();
    u_points = nullptr;
    u_sort = nullptr;
    tree = RadixTreeData{};
    u_edge_count = nullptr;
    u_prefix_sum = nullptr;
  }

  int num_unique = 0;
  int num_oct_nodes = 0;
  float* u_points = nullptr;  // n float4
  unsigned int* u_sort = nullptr;
  RadixTreeData tree;
  int* u_edge_count = nullptr;
  int* u_prefix_sum = nullptr;  // tree.n_nodes + 1 entries
  ocp_ctx ctx = nullptr;        // the context every stage ran on

  int n() const { return n_; }

 private:
  void check(int rc, const char* what) {
    if (rc != OCP_OK)
      throw std::runtime_error(std::string(what) + ": status " + std::to_string(rc) + ": " +
                               (ctx ? ocp_last_error(ctx) : "no context"));
  }
  template <class T>
  T* alloc(size_t count) {
    void* p = nullptr;
    check(ocp_malloc(ctx, &p, std::max<size_t>(count, 1) * sizeof(T)), "ocp_malloc");
    owned_.push_back(p);
    return static_cast<T*>(p);
  }

  int n_;
  int device_;
  std::vector<void*> owned_;
};

}  // namespace ocp

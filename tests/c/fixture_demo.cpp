// A C++ caller of include/octree_fixture.hpp: sets up ocp::GpuFixture as the
// reference's GPU benchmarks do (benchmarks/cuda/fixture.cuh:26-69), runs the
// octree stage on its members as bm_07_octree.cuh drives k_MakeOctNodes and
// k_LinkLeafNodes, then checks every array against one fused ocp_build of the
// same points: unique keys, radix tree, edge counts, offsets and nodes must
// match byte for byte. Prints "U K" on success. Without a GPU it exits 3.
//
//   fixture_demo [n] [points_out]   (points_out: the n float4 points, raw)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "octree_fixture.hpp"

namespace {

template <class T>
std::vector<T> host_copy(ocp_ctx ctx, const T* dev, size_t count) {
  std::vector<T> h(count);
  if (count && ocp_memcpy_d2h(ctx, h.data(), dev, count * sizeof(T)) != OCP_OK)
    throw std::runtime_error(std::string("copy: ") + ocp_last_error(ctx));
  return h;
}

template <class T>
bool same(ocp_ctx a, const T* da, ocp_ctx b, const T* db, size_t count, const char* what) {
  const std::vector<T> x = host_copy(a, da, count), y = host_copy(b, db, count);
  if (count && std::memcmp(x.data(), y.data(), count * sizeof(T)) != 0) {
    std::fprintf(stderr, "mismatch: %s\n", what);
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : ocp::kN;
  ocp_ctx probe = nullptr;
  if (ocp_create(&probe, 0, nullptr) != OCP_OK) {
    std::fprintf(stderr, "no device\n");
    return 3;
  }
  ocp_destroy(probe);
  try {
    ocp::GpuFixture fx(n);
    fx.SetUp();
    if (argc > 2) {
      const std::vector<float> pts = host_copy(fx.ctx, fx.u_points, size_t(n) * 4);
      FILE* f = std::fopen(argv[2], "wb");
      if (!f || std::fwrite(pts.data(), sizeof(float), pts.size(), f) != pts.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
      }
      std::fclose(f);
    }
    // The octree stage on the fixture's arrays (bm_07_octree.cuh).
    void* dn = nullptr;
    if (ocp_malloc(fx.ctx, &dn, size_t(fx.num_oct_nodes) * sizeof(ocp_oct_node)) != OCP_OK)
      throw std::runtime_error("node buffer");
    auto* nodes = static_cast<ocp_oct_node*>(dn);
    const int m = fx.tree.n_nodes;
    int rc = ocp_k_InitOctNodes(fx.ctx, nodes, fx.num_oct_nodes, fx.u_sort, fx.tree.prefixN, m,
                                ocp::kMin, ocp::kRange);
    if (!rc)
      rc = ocp_k_MakeOctNodes(fx.ctx, nodes, fx.u_prefix_sum, fx.u_edge_count, fx.u_sort,
                              fx.tree.prefixN, fx.tree.parent, ocp::kMin, ocp::kRange, m,
                              fx.num_oct_nodes);
    if (!rc)
      rc = ocp_k_LinkLeafNodes(fx.ctx, nodes, fx.u_prefix_sum, fx.u_edge_count, fx.u_sort,
                               fx.tree.hasLeafLeft, fx.tree.hasLeafRight, fx.tree.prefixN,
                               fx.tree.parent, fx.tree.leftChild, m, fx.num_oct_nodes);
    if (!rc) rc = ocp_synchronize(fx.ctx);
    if (rc) throw std::runtime_error(std::string("octree stage: ") + ocp_last_error(fx.ctx));

    // One fused build of the same points.
    ocp_ctx c2 = nullptr;
    if (ocp_create(&c2, 0, nullptr) != OCP_OK) throw std::runtime_error("second context");
    ocp_results r;
    int32_t u = 0, k = 0;
    rc = ocp_build(c2, fx.u_points, 0, n, ocp::kMin, ocp::kRange);
    if (!rc) rc = ocp_get_results(c2, &r);
    if (!rc) rc = ocp_get_counts(c2, &u, &k);
    if (rc) throw std::runtime_error(std::string("build: ") + ocp_last_error(c2));
    bool ok = u == fx.num_unique && k == fx.num_oct_nodes;
    if (!ok) std::fprintf(stderr, "counts: fixture %d %d, build %d %d\n", fx.num_unique,
                          fx.num_oct_nodes, u, k);
    const size_t um = size_t(m);
    ok = ok && same(fx.ctx, fx.u_sort, c2, r.unique_keys, size_t(u), "unique keys");
    ok = ok && same(fx.ctx, fx.tree.prefixN, c2, r.prefix_n, um, "prefixN");
    ok = ok && same(fx.ctx, fx.tree.hasLeafLeft, c2, r.has_leaf_left, um, "hasLeafLeft");
    ok = ok && same(fx.ctx, fx.tree.hasLeafRight, c2, r.has_leaf_right, um, "hasLeafRight");
    ok = ok && same(fx.ctx, fx.tree.leftChild, c2, r.left_child, um, "leftChild");
    ok = ok && same(fx.ctx, fx.tree.parent, c2, r.parent, um, "parent");
    ok = ok && same(fx.ctx, fx.u_edge_count, c2, r.edge_count, um, "edge counts");
    ok = ok && same(fx.ctx, fx.u_prefix_sum, c2, r.offsets, um, "offsets");
    ok = ok && same(fx.ctx, static_cast<const ocp_oct_node*>(nodes), c2, r.nodes,
                    size_t(k), "nodes");
    ocp_destroy(c2);
    ocp_free(fx.ctx, dn);
    if (!ok) return 1;
    std::printf("%d %d\n", fx.num_unique, fx.num_oct_nodes);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}

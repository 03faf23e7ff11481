/*
 * C-ABI wrappers around the reference's own CPU stage functions, compiled
 * together with the reference sources that need no third-party header
 * (kernels/02_sort.cpp, 03_unique.cpp, 04_radix_tree.cpp, 05_edge_count.cpp,
 * 06_partial_sum.cpp) into oracle/_ref/libocref.so by oracle/Makefile.
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle and to write golden
 * fixtures in the build container. Never shipped, never on the product path.
 *
 * The declarations below restate the reference prototypes from
 * include/kernels/02_sort.hpp:5-7, 03_unique.hpp:4, 04_radix_tree.hpp:5-11,
 * 05_edge_count.hpp:5-8 and 06_partial_sum.hpp:6-9 so that the wrappers link
 * against the reference objects.
 */
#include <omp.h>

#include <cstddef>
#include <cstdint>
#include <cstring>

void k_SortKeysInplace(unsigned int* keys, int n);
void k_SimpleRadixSort(unsigned int* keys, int n);
int k_CountUnique(unsigned int* keys, int n);
void k_BuildRadixTree(int n, const unsigned int* codes, uint8_t* prefix_n,
                      bool* has_leaf_left, bool* has_leaf_right,
                      int* left_child, int* parent);
void k_EdgeCount(const uint8_t* prefix_n, const int* parents, int* edge_count,
                 int n_brt_nodes);
int* k_PartialSum(const int* data, std::ptrdiff_t first, std::ptrdiff_t last,
                  int* d_first);

extern "C" {

void ref_simple_radix_sort(uint32_t* keys, int n, int nthreads) {
  omp_set_num_threads(nthreads);
  k_SimpleRadixSort(keys, n);
}

int ref_count_unique(uint32_t* keys, int n) { return k_CountUnique(keys, n); }

/* Zero-filled outputs, one thread (ascending i), n = number of BRT nodes.
 * The caller keeps n >= 2 (n == 1 reaches clz(0) in the reference). */
void ref_build_radix_tree(int n, const uint32_t* codes, uint8_t* prefix_n,
                          uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                          int32_t* left_child, int32_t* parent) {
  std::memset(prefix_n, 0, n);
  std::memset(has_leaf_left, 0, n);
  std::memset(has_leaf_right, 0, n);
  std::memset(left_child, 0, sizeof(int32_t) * n);
  std::memset(parent, 0, sizeof(int32_t) * n);
  omp_set_num_threads(1);
  k_BuildRadixTree(n, codes, prefix_n, reinterpret_cast<bool*>(has_leaf_left),
                   reinterpret_cast<bool*>(has_leaf_right), left_child, parent);
}

void ref_edge_count(const uint8_t* prefix_n, const int32_t* parent,
                    int32_t* edge_count, int n_brt) {
  omp_set_num_threads(1);
  k_EdgeCount(prefix_n, parent, edge_count, n_brt);
}

void ref_partial_sum(const int32_t* in, int n, int32_t* out) {
  (void)k_PartialSum(in, 0, n, out);
}

}  // extern "C"

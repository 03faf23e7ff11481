/*
 * CPU oracle for the seven-stage octree pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py). Nothing in the product path (pipeline_amd/) links or calls it.
 *
 * It restates the reference CPU path (the kernels/0N_*.cpp files and the
 * include/shared/ headers of
 * ucsc-redwood/pipeline) as plain C-ABI functions, run single-threaded in
 * ascending index order on zero-initialised buffers (the canonical semantics),
 * plus OpenMP variants with the reference's own loop structure that are used
 * only for timing.
 */
#ifndef OCTREE_ORACLE_H
#define OCTREE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 60-byte node, layout of OctNode in include/shared/types.h:28-34
 * (glm::vec4 with its default 4-byte alignment). */
typedef struct orc_oct_node {
  int32_t children[8];
  float corner[4];
  float cell_size;
  int32_t child_node_mask;
  int32_t child_leaf_mask;
} orc_oct_node;

/* ---- scalar helpers (exported so tests can pin them directly) ---------- */
uint32_t orc_split_by_3(uint32_t a);                       /* morton.h:22-29 */
uint32_t orc_encode(float x, float y, float z, float min_coord, float range); /* morton.h:38-46 */
uint32_t orc_third_bits(uint32_t m);                       /* morton.h:52-59 */
void orc_decode(uint32_t code, float min_coord, float range, float out[4]); /* morton.h:67-87 */
uint32_t orc_delta(uint32_t a, uint32_t b);                /* 04_radix_tree.cpp:21-27 */
int32_t orc_log2_ceil(uint32_t x);                         /* 04_radix_tree.cpp:29-37 */

/* ---- stages -------------------------------------------------------------- */
/* Stage 1, kernels/01_morton.cpp:5-14. pts is n x float4 (w ignored). */
void orc_morton(const float* pts, uint32_t* keys, int64_t n, float min_coord,
                float range, int nthreads);
/* Stage 1 with a 12-byte float3 input (declared variant). */
void orc_morton3(const float* pts, uint32_t* keys, int64_t n, float min_coord,
                 float range);
/* Stage 2, kernels/02_sort.cpp:13-58 (LSD, 8-bit digits, 4 passes). */
void orc_radix_sort(uint32_t* keys, int64_t n, int nthreads);
/* Stage 3, kernels/03_unique.cpp:2-16 (in place). Returns U. */
int32_t orc_count_unique(uint32_t* keys, int32_t n);
/* Stage 4, kernels/04_radix_tree.cpp:39-121. n = number of BRT nodes (U-1).
 * Outputs are zero-filled first; parent[0] stays 0. */
void orc_build_radix_tree(int32_t n, const uint32_t* codes, uint8_t* prefix_n,
                          uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                          int32_t* left_child, int32_t* parent, int nthreads);
/* Stage 5, kernels/05_edge_count.cpp:3-16, over n_brt nodes. */
void orc_edge_count(const uint8_t* prefix_n, const int32_t* parent,
                    int32_t* edge_count, int32_t n_brt, int nthreads);
/* Stage 6, kernels/06_partial_sum.cpp:3-15 (inclusive scan). */
void orc_partial_sum(const int32_t* in, int32_t n, int32_t* out);
/* Number of octree node slots the build touches: max(1, max_i P[i]+e[i]);
 * 0 when U == 0. */
int32_t orc_oct_alloc_count(const int32_t* offsets, const int32_t* counts,
                            int32_t n_brt, int32_t n_unique);
/* Root node, apps/cpu_demo/main.cpp:125-132. */
void orc_root_init(orc_oct_node* nodes, const uint32_t* codes,
                   const uint8_t* prefix_n, int32_t n_brt, float min_coord,
                   float range);
/* Stage 7a, kernels/07_octree.cpp:5-31 -> include/shared/oct.h:24-90.
 * n_alloc bounds every slot index written. */
void orc_make_oct_nodes(orc_oct_node* nodes, const int32_t* offsets,
                        const int32_t* counts, const uint32_t* codes,
                        const uint8_t* prefix_n, const int32_t* parent,
                        float min_coord, float range, int32_t n_brt,
                        int32_t n_alloc, int nthreads);
/* Stage 7b, kernels/07_octree.cpp:33-57 -> include/shared/oct.h:92-134. */
void orc_link_leaf_nodes(orc_oct_node* nodes, const int32_t* offsets,
                         const int32_t* counts, const uint32_t* codes,
                         const uint8_t* has_leaf_left,
                         const uint8_t* has_leaf_right,
                         const uint8_t* prefix_n, const int32_t* parent,
                         const int32_t* left_child, int32_t n_brt,
                         int32_t n_alloc, int nthreads);

/* Whole chain in the order of apps/cpu_demo/main.cpp:63-155, timed per stage
 * with OpenMP on nthreads host threads. Fills stage_ms[8] with the median of
 * reps runs (morton, sort, unique, tree, edge, scan, octree, total) and
 * returns U. */
int32_t orc_time_pipeline(const float* pts, int64_t n, float min_coord,
                          float range, int nthreads, int reps,
                          double* stage_ms, int32_t* out_k);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif

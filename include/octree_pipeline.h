/*
 * octree_pipeline.h - C ABI of the MI355X (gfx950) octree-construction
 * pipeline: Morton encode, OneSweep radix sort, unique, Karras radix tree,
 * edge count, prefix sum, octree nodes.
 *
 * Conventions
 *  - Every function returns an ocp_status; 0 is success. The text of the last
 *    failure is ocp_last_error(ctx). Nothing exits or throws across the ABI.
 *  - All array arguments are device pointers (allocate with ocp_malloc or any
 *    hipMalloc'd memory of the same process). Counts that the reference
 *    passes by value are host integers.
 *  - Work is enqueued asynchronously on the context's stream; ocp_synchronize
 *    waits for it. One context per host thread; contexts are not locked.
 *  - Layouts are the reference's: float4 points (16 B, w ignored) or the
 *    declared float3 variant (12 B), uint32 keys, uint8 prefix lengths,
 *    uint8 booleans, int32 indices, 60-byte OctNode.
 *  - Results are bit-identical to the reference CPU path (kernels/0N_*.cpp of
 *    ucsc-redwood/pipeline) run single-threaded in ascending index order on
 *    zeroed buffers.
 *
 * Per-stage entry points replace the reference's `gpu::k_*` kernels
 * (include/cuda/kernels/0N_*.cuh) and CPU stages (include/kernels/0N_*.hpp);
 * each declaration cites the interface it replaces.
 */
#ifndef OCTREE_PIPELINE_H
#define OCTREE_PIPELINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocp_context* ocp_ctx;

typedef enum ocp_status {
  OCP_OK = 0,
  OCP_ERR_INVALID = 1,   /* bad argument: null pointer, size, table */
  OCP_ERR_HIP = 2,       /* a HIP runtime call failed */
  OCP_ERR_CAPACITY = 3,  /* an input exceeds a documented maximum */
  OCP_ERR_STATE = 4      /* no result yet, or the call sequence is wrong */
} ocp_status;

/* include/shared/types.h:28-34 (60 bytes, 4-byte aligned). */
typedef struct ocp_oct_node {
  int32_t children[8];
  float corner[4];
  float cell_size;
  int32_t child_node_mask;
  int32_t child_leaf_mask;
} ocp_oct_node;

/* ------------------------------------------------------------ context */
/* stream: a hipStream_t, or NULL for the context's own stream. */
int ocp_create(ocp_ctx* out, int device, void* stream);
int ocp_destroy(ocp_ctx ctx);
const char* ocp_last_error(ocp_ctx ctx);
/* Waits for the context's stream. Returns OCP_ERR_STATE when a lookback
 * wait of the work enqueued since the last check gave up (see
 * ocp_set_wait_limit): that work's output is invalid, the pending build
 * result is dropped, and the condition is cleared, so the next call runs
 * normally. ocp_get_results, ocp_get_counts and the synchronous entries
 * report it the same way. The reference checks with checkCudaErrors and
 * exits instead (include/cuda/common/helper_cuda.hpp:583-602). */
int ocp_synchronize(ocp_ctx ctx);
/* Every wait for an earlier tile (the decoupled lookbacks of the sort,
 * unique and scan stages) gives up after `polls` unsuccessful polls
 * (polls < 0: the default, 2^22). Once one has given up, every later
 * lookback-stage kernel on the context does nothing until the condition is
 * reported. With polls = 0 every wait gives up at once and every kernel
 * still runs to its end: a way to exercise the error path. No counterpart
 * in the reference, whose waits spin unbounded
 * (kernels/cuda/02_sort.cu:688-715, 735-755). */
int ocp_set_wait_limit(ocp_ctx ctx, int64_t polls);
/* Plumbing for callers without their own allocator (tests, bindings). */
int ocp_malloc(ocp_ctx ctx, void** ptr, size_t bytes);
int ocp_free(ocp_ctx ctx, void* ptr);
int ocp_memcpy_h2d(ocp_ctx ctx, void* dst, const void* src, size_t bytes);
int ocp_memcpy_d2h(ocp_ctx ctx, void* dst, const void* src, size_t bytes);
int ocp_memset(ocp_ctx ctx, void* dst, int value, size_t bytes);
/* Free and total bytes of the context's device (hipMemGetInfo), e.g. to
 * size a batch against ocp_workspace_size. */
int ocp_device_memory(ocp_ctx ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Device-to-device copy kernel on the context stream (16-byte aligned
 * pointers, size a multiple of 16). */
int ocp_copy(ocp_ctx ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------ input */
/* Replaces gpu::k_InitRandomVec4 (include/cuda/kernels/00_init.cuh:7-8,
 * kernels/cuda/00_init.cu:27-50): n float4 points from the hybrid
 * LCG-Tausworthe generator, one step per point from state seeded by its
 * index - the reference's output when launched with at least n threads
 * (apps/gpu_demo/main.cu:167-173). Coordinates (and w) in [min, min+range). */
int ocp_k_InitRandomVec4(ocp_ctx ctx, float* data, int32_t n, float min_coord,
                         float range, int32_t seed);

/* ------------------------------------------------------------ stage 1 */
/* Replaces gpu::k_ComputeMortonCode (include/cuda/kernels/01_morton.cuh:7-11)
 * and k_ComputeMortonCode (include/kernels/01_morton.hpp:9-13).
 * data: n float4, 16-byte aligned. */
int ocp_k_ComputeMortonCode(ocp_ctx ctx, const float* data,
                            uint32_t* morton_keys, size_t n, float min_coord,
                            float range);
/* Same with a packed float3 input (12 bytes per point). */
int ocp_k_ComputeMortonCode3(ocp_ctx ctx, const float* data,
                             uint32_t* morton_keys, size_t n, float min_coord,
                             float range);

/* ------------------------------------------------------------ stage 2 */
/* Replaces gpu::k_GlobalHistogram_WithLogicalBlocks
 * (include/cuda/kernels/02_sort.cuh:21-24): raw 4 x 256 digit counts of
 * `size` keys into global_histogram[1024] (zeroed by the call). */
int ocp_k_GlobalHistogram(ocp_ctx ctx, const uint32_t* sort,
                          uint32_t* global_histogram, uint32_t size);
/* Replaces gpu::k_DigitBinning_WithLogicalBlocks
 * (include/cuda/kernels/02_sort.cuh:25-33): one stable OneSweep pass on the
 * digit at radix_shift (0, 8, 16, 24) from `sort` into `alt`, using the
 * counts of ocp_k_GlobalHistogram. The lookback scratch and tile counter
 * the reference takes as passHistogram/index live in the context and need
 * no reset between calls. */
int ocp_k_DigitBinning(ocp_ctx ctx, const uint32_t* global_histogram,
                       const uint32_t* sort, uint32_t* alt, uint32_t size,
                       uint32_t radix_shift);
/* Replaces k_SimpleRadixSort(keys, keys_alt, n)
 * (include/kernels/02_sort.hpp:7) and the reference's DispatchSortKernels
 * (apps/gpu_demo/main.cu:63-125): histogram + 4 passes
 * keys -> alt -> keys -> alt -> keys; the sorted keys end in `keys`. */
int ocp_k_SortKeys(ocp_ctx ctx, uint32_t* keys, uint32_t* keys_alt,
                   uint32_t n);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ocp_k_SortKeys with every scratch buffer (digit counts,
 * lookback flags and their epoch) in a fixed array of the library's device
 * memory, 8 MiB per device, set up once per device by ocp_create. Makes no
 * allocation, frees nothing. At most 2^24 keys (OCP_ERR_CAPACITY beyond).
 * All calls on one device share that array and keep its last contents:
 * calls that overlap on one device, from any context or stream, race. */
#define OCP_STATIC_SORT_MAX_KEYS (1u << 24)
int ocp_k_SortKeys_static(ocp_ctx ctx, uint32_t* keys, uint32_t* keys_alt,
                          uint32_t n);

/* The key->point permutation, which the reference does not keep (it sorts
 * bare keys, kernels/cuda/02_sort.cu:566-574, and uniques in place,
 * kernels/03_unique.cpp:2-16): a stable sort of n keys carrying one 32-bit
 * value each (histogram + 4 stable passes; keys/vals -> alt -> keys/vals).
 * With vals = 0..n-1 the values end as the stable argsort of the keys. */
int ocp_k_SortPairs(ocp_ctx ctx, uint32_t* keys, uint32_t* vals,
                    uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n);

/* ------------------------------------------------------------ stage 3 */
/* Replaces gpu::k_CountUnique (include/cuda/kernels/03_unique.cuh:5) and
 * k_CountUnique (include/kernels/03_unique.hpp:4): in-place unique of n
 * sorted keys; the count is written to num_unique_out (device int32). */
int ocp_k_CountUnique(ocp_ctx ctx, uint32_t* keys, int32_t* num_unique_out,
                      int32_t n);
/* The same with a value per key: each kept key keeps the value of the first
 * key of its run (after ocp_k_SortPairs, its smallest point index). */
int ocp_k_CountUniquePairs(ocp_ctx ctx, uint32_t* keys, uint32_t* vals,
                           int32_t* num_unique_out, int32_t n);

/* ------------------------------------------------------------ stage 4 */
/* Replaces gpu::k_BuildRadixTree (include/cuda/kernels/04_radix_tree.cuh:7-13)
 * and k_BuildRadixTree (include/kernels/04_radix_tree.hpp:5-11): n nodes
 * over n+1 sorted unique codes; pass n = U-1. parent[0] is written 0. */
int ocp_k_BuildRadixTree(ocp_ctx ctx, int32_t n, const uint32_t* codes,
                         uint8_t* prefix_n, uint8_t* has_leaf_left,
                         uint8_t* has_leaf_right, int32_t* left_child,
                         int32_t* parent);

/* ------------------------------------------------------------ stage 5 */
/* Replaces gpu::k_EdgeCount (include/cuda/kernels/05_edge_count.cuh:7-10)
 * and k_EdgeCount (include/kernels/05_edge_count.hpp:5-8). */
int ocp_k_EdgeCount(ocp_ctx ctx, const uint8_t* prefix_n,
                    const int32_t* parents, int32_t* edge_count,
                    int32_t n_brt_nodes);

/* ------------------------------------------------------------ stage 6 */
/* Replaces k_PartialSum (include/kernels/06_partial_sum.hpp:6-9), which the
 * reference GPU path runs on the host (apps/gpu_demo/main.cu:308-310):
 * inclusive scan of data[first, last) into d_first. */
int ocp_k_PartialSum(ocp_ctx ctx, const int32_t* data, int64_t first,
                     int64_t last, int32_t* d_first);

/* ------------------------------------------------------------ stage 7 */
/* The host-side root setup of apps/cpu_demo/main.cpp:120-132 on device:
 * zero n_alloc nodes, then node 0 = corner of codes[0]'s root-level prefix
 * (root level = prefix_n[0]/3, 0 when n_brt == 0), cell_size = range. */
int ocp_k_InitOctNodes(ocp_ctx ctx, ocp_oct_node* nodes, int32_t n_alloc,
                       const uint32_t* codes, const uint8_t* prefix_n,
                       int32_t n_brt, float min_coord, float range);
/* Replaces gpu::k_MakeOctNodes (include/cuda/kernels/07_octree.cuh:7-17) and
 * k_MakeOctNodes (include/kernels/07_octree.hpp:5-14). N = number of BRT
 * nodes; n_alloc = number of node slots (max(1, max_i P[i]+e[i])), which
 * bounds every write. Expects nodes prepared by ocp_k_InitOctNodes. */
int ocp_k_MakeOctNodes(ocp_ctx ctx, ocp_oct_node* oct_nodes,
                       const int32_t* node_offsets,
                       const int32_t* rt_node_counts, const uint32_t* codes,
                       const uint8_t* rt_prefixN, const int32_t* rt_parents,
                       float min_coord, float range, int32_t N,
                       int32_t n_alloc);
/* Replaces gpu::k_LinkLeafNodes (include/cuda/kernels/07_octree.cuh:19-28)
 * and k_LinkLeafNodes (include/kernels/07_octree.hpp:16-25). */
int ocp_k_LinkLeafNodes(ocp_ctx ctx, ocp_oct_node* nodes,
                        const int32_t* node_offsets,
                        const int32_t* rt_node_counts, const uint32_t* codes,
                        const uint8_t* rt_hasLeafLeft,
                        const uint8_t* rt_hasLeafRight,
                        const uint8_t* rt_prefixN, const int32_t* rt_parents,
                        const int32_t* rt_leftChild, int32_t N,
                        int32_t n_alloc);

/* ------------------------------------------------------------ pipeline */
/* The whole chain of apps/gpu_demo/main.cu:157-356 / benchmarks/cuda/
 * fixture.cuh:26-69 on device, with no host round trip: stage outputs stay
 * in context-owned buffers (ocp_results) and the counts U and K are written
 * to device memory. points: n float4 (float3 != 0: packed float3).
 * Replays a captured hipGraph when graph mode is on. */
int ocp_build(ocp_ctx ctx, const float* points, int float3, int64_t n,
              float min_coord, float range);
/* Batched form: n_seg independent clouds, cloud k = points
 * [seg_offsets[k], seg_offsets[k+1]) with its own (seg_min[k],
 * seg_range[k]); one segmented launch per stage. Host arrays. With seg_min
 * and seg_range both NULL, each cloud's cube box is reduced on the device
 * (min over x, y, z; range = max((max - min) * 1.0001, 1e-3), so the far
 * face quantises below 1024) and read back through ocp_results. */
int ocp_build_batched(ocp_ctx ctx, const float* points, int float3,
                      const int32_t* seg_offsets, const float* seg_min,
                      const float* seg_range, int32_t n_seg);
/* The checks ocp_build_batched applies to its table, without a context or
 * device: offsets start at 0 and do not decrease, each segment holds at most
 * 2^27 points, the batch at most 2^30-1, ranges are > 0 (boxes may be both
 * NULL). */
int ocp_check_batch(const int32_t* seg_offsets, const float* seg_min,
                    const float* seg_range, int32_t n_seg);
/* Device memory a context allocates, without a context or device: for a
 * first build of n points in n_seg clouds (n_seg 0: plain) with this tile
 * table (NULL: the defaults) and the permutation on or off, and the scratch
 * of the per-stage sort, unique and scan entries on n entries (the
 * reference's OneSweepData and friends, which its caller allocates,
 * include/shared/types.h:42-65). A build can grow the node buffer later if
 * its octree needs more than n slots. */
typedef struct ocp_workspace_sizes {
  int64_t build_bytes;   /* ocp_build / ocp_build_batched workspace */
  int64_t sort_bytes;    /* ocp_k_SortKeys, ocp_k_DigitBinning, ocp_k_SortPairs */
  int64_t unique_bytes;  /* ocp_k_CountUnique, ocp_k_CountUniquePairs */
  int64_t scan_bytes;    /* ocp_k_PartialSum, ocp_k_EdgeCount */
} ocp_workspace_sizes;
struct ocp_tune;
int ocp_workspace_size(const struct ocp_tune* tune, int64_t n, int32_t n_seg,
                       int permutation, ocp_workspace_sizes* out);
/* ocp_build with the cloud's cube box reduced on the device (as for
 * ocp_build_batched with NULL boxes): no host value is needed before the
 * build. The reference fixes the box at [0, 1024) (benchmarks/config.hpp:
 * 7-10). */
int ocp_build_autobox(ocp_ctx ctx, const float* points, int float3, int64_t n);
/* Moves a cloud in device memory: p' = clamp(A p + t, lo, hi) per coordinate,
 * with affine = {A row-major (9 floats), t (3 floats)}; a float4 point's
 * fourth value passes through; src and dst are the same buffer or do not
 * overlap. Asynchronous on the context stream. Successive frames of a
 * moving sensor are made this way (rigid motion: A a rotation). No
 * counterpart in the reference, whose frames are files on disk. */
int ocp_transform_points(ocp_ctx ctx, const float* src, float* dst, int float3, int64_t n,
                         const float* affine, float lo, float hi);
/* Builds carry the key->point permutation when enabled: every sort pass
 * moves a point index with its key (and is stable), unique keeps each key's
 * first point, and ocp_results gives both (sorted_index, unique_index).
 * Indices are relative to each cloud's first point. Off by default. */
int ocp_set_permutation(ocp_ctx ctx, int enable);
/* Capture each distinct build shape once into a hipGraph and replay it. */
int ocp_set_graph_mode(ocp_ctx ctx, int enable);

/* Device pointers to the last build's outputs, indexed by point offset of
 * each segment (node arrays by oct_start). */
typedef struct ocp_results {
  const uint32_t* sorted_keys;    /* n */
  const uint32_t* unique_keys;    /* first U_k entries of each segment */
  const uint8_t* prefix_n;        /* first U_k - 1 entries of each segment */
  const uint8_t* has_leaf_left;
  const uint8_t* has_leaf_right;
  const int32_t* left_child;
  const int32_t* parent;
  const int32_t* edge_count;
  const int32_t* offsets;         /* inclusive scan of edge_count */
  const ocp_oct_node* nodes;      /* segment k at oct_start[k] */
  const int32_t* n_unique;        /* device, n_seg */
  const int32_t* n_oct;           /* device, n_seg: K_k */
  const int32_t* oct_start;       /* device, n_seg + 1 */
  int32_t n_seg;
  int64_t n_points;
  const float* seg_min;           /* device, n_seg: the boxes a batched or */
  const float* seg_range;         /* autobox build used (NULL otherwise)   */
  const uint32_t* sorted_index;   /* n: point of each sorted key, and the  */
  const uint32_t* unique_index;   /* first point of each unique key (NULL  */
                                  /* unless ocp_set_permutation is on)     */
} ocp_results;
int ocp_get_results(ocp_ctx ctx, ocp_results* out);
/* Synchronises and copies the per-segment U and K (host arrays of n_seg). */
int ocp_get_counts(ocp_ctx ctx, int32_t* n_unique, int32_t* n_oct);

/* ------------------------------------------------------------ streaming
 * The pipelined demo `ppl` that apps/xmake.lua:7 and apps/README.md:5 name
 * (its sources are not in the reference): a stream of frames, each built
 * as by ocp_build, with consecutive frames overlapping on the device.
 * Frames rotate over `depth` slots; each slot is a context with its own
 * stream and outputs, a pinned host buffer and a device points buffer of
 * max_points float4. A slot's upload and build run in order on its stream,
 * so one frame's upload and Morton/radix-tree/octree kernels run beside
 * another frame's work; the sort, unique and scan launches of two frames
 * are ordered (their tiles wait on each other inside a launch). Slot
 * contexts run without graphs. */
typedef struct ocp_pipeline* ocp_ppl;
int ocp_ppl_create(ocp_ppl* out, int device, int32_t depth, int64_t max_points);
int ocp_ppl_destroy(ocp_ppl ppl);
const char* ocp_ppl_last_error(ocp_ppl ppl);
/* The slot's pinned host buffer (max_points float4, or float3 packed), for
 * the producer to write the next frame into; waits until the slot's last
 * upload from it has finished. */
int ocp_ppl_host_buffer(ocp_ppl ppl, int32_t slot, float** host);
/* Enqueues one frame on a slot: points is a device array, or NULL to upload
 * n points from the slot's host buffer first. The slot's previous outputs
 * are overwritten: read them (ocp_ppl_wait, then ocp_get_results on
 * ocp_ppl_context) before submitting to the slot again. Asynchronous. */
int ocp_ppl_submit(ocp_ppl ppl, int32_t slot, const float* points, int float3,
                   int64_t n, float min_coord, float range);
/* Waits for the slot's frame; reports its errors as ocp_synchronize does. */
int ocp_ppl_wait(ocp_ppl ppl, int32_t slot);
/* The slot's context: its results, validator and compact form. */
int ocp_ppl_context(ocp_ppl ppl, int32_t slot, ocp_ctx* ctx);

/* ------------------------------------------------------------ octree checks
 * Structural check of the last build's node array on the device, in the role
 * of checkTree (apps/cpu_demo/main.cpp:13-35, which the reference calls on
 * its CPU output). The reference's offsets are an inclusive scan used as
 * first slots (kernels/06_partial_sum.cpp:3-15, include/shared/oct.h:24-90),
 * so chains overlap and child-in-octant does not hold in its own output; the
 * counted violations are the properties that hold whatever the write order.
 * The last two fields are information, not violations. Synchronises. */
typedef struct ocp_tree_report {
  int64_t slots;              /* K summed over segments */
  int64_t written;            /* slots with a nonzero byte */
  int64_t covered_unwritten;  /* violation: slot 0 or a chain slot left zero */
  int64_t written_uncovered;  /* violation: written outside every chain */
  int64_t chain_past_end;     /* violation: chain [P, P + e) not inside K */
  int64_t bad_w;              /* violation: corner w != 1 */
  int64_t bad_cell;           /* violation: cell size not range / 2^k */
  int64_t off_grid;           /* violation: corner off its level's grid */
  int64_t leaf_mask;          /* violation: child_leaf_mask != 0 */
  int64_t bad_root;           /* violation: root corner or cell size */
  int64_t links;              /* node-mask links */
  int64_t links_in_octant;    /* of those, child inside the parent's octant */
} ocp_tree_report;
int ocp_validate_octree(ocp_ctx ctx, ocp_tree_report* report);

/* Compact node format (12 B header per slot plus the nonzero children in one
 * array), an exact encoding of the 60-byte OctNode of include/shared/
 * types.h:28-34: the corner is the decode of a Morton prefix
 * (include/shared/morton.h:45-80) and the cell size range / 2^depth
 * (include/shared/oct.h:56-57, 87-88), so both are stored as integers. The
 * reference has no compact form; this is its node array at a fifth of the
 * bytes or less, for storage and transfer. */
typedef struct ocp_compact_node {
  uint32_t prefix;       /* Morton code of the corner (level's prefix) */
  uint8_t level;         /* octree level | depth below the root << 4;
                            0xff: slot never written (all zero) */
  uint8_t node_mask;     /* child_node_mask */
  uint8_t child_set;     /* bit c: children[c] != 0 */
  uint8_t leaf_mask;     /* child_leaf_mask */
  uint32_t first_child;  /* index of its first stored child */
} ocp_compact_node;
/* Compacts the last build's nodes (all segments, in oct_start order) into
 * device arrays. With nodes or children NULL, only writes the sizes
 * (n_nodes = K total, n_children). OCP_ERR_CAPACITY when a cap is below
 * them; OCP_ERR_STATE when a slot has no exact compact form (not the case
 * for the pipeline's own output). Synchronises. */
int ocp_compact_octree(ocp_ctx ctx, ocp_compact_node* nodes, int64_t node_cap,
                       int32_t* children, int64_t child_cap, int64_t* n_nodes,
                       int64_t* n_children);
/* 60-byte nodes from the compact form: oct_start (device, n_seg + 1) lays
 * the segments out as in ocp_results; boxes are per segment (seg_min,
 * seg_range device arrays) or, both NULL, min_coord and range. Child indices
 * past n_children read as 0. Asynchronous. */
int ocp_expand_octree(ocp_ctx ctx, const ocp_compact_node* nodes,
                      const int32_t* children, int64_t n_children,
                      const int32_t* oct_start, int32_t n_seg,
                      const float* seg_min, const float* seg_range,
                      float min_coord, float range, ocp_oct_node* out);

/* Per-stage device time of the last eager build, in ms, measured with HIP
 * events on the context stream when stage timing is on: [morton, sort,
 * unique, radix_tree, edge_scan, octree, total]. */
int ocp_set_stage_timing(ocp_ctx ctx, int enable);
/* Cold per-kernel timing: with flush_bytes > 0 (above the 256 MB Infinity
 * Cache, e.g. 512 MiB), each kernel timed by stage timing is preceded by a
 * fill of a context-owned buffer of that size, outside its events - the
 * role of CudaEventTimer's L2 flush (benchmarks/cuda/
 * cuda_bench_helper.cuh:69-90). 0 frees the buffer. Measurement only. */
int ocp_set_cold_timing(ocp_ctx ctx, int64_t flush_bytes);
int ocp_get_stage_ms(ocp_ctx ctx, float* ms7);
/* Per-kernel device time of the last eager build with stage timing on, in
 * launch order, each bracketed by HIP events on the context stream. Writes
 * up to max_kernels entries (names are static strings) and the number of
 * kernels a build launches to n_kernels. A build with the MSD sort
 * (ocp_tune.sort_msd) names its sort slots msd_scatter, msd_buckets,
 * msd_lsd_fallback (msd_lsd_skipped when every bucket fit in LDS) and
 * unused. */
int ocp_get_kernel_ms(ocp_ctx ctx, float* ms, const char** names,
                      int32_t max_kernels, int32_t* n_kernels);
/* Time `steps` builds (after `warmup`) of the configured kind on the
 * context stream with HIP events, bracketed by device synchronisation.
 * Writes the mean ms per build. */
int ocp_time_builds(ocp_ctx ctx, const float* points, int float3,
                    const int32_t* seg_offsets, const float* seg_min,
                    const float* seg_range, int32_t n_seg, int warmup,
                    int steps, float* ms_per_build);

/* ------------------------------------------------------------ tuning */
/* Tile parameters and the sort's digit width (pass count). */
typedef struct ocp_tune {
  int32_t sort_waves;   /* OneSweep tile of 8-bit digits: waves ... */
  int32_t sort_kpt;     /* ... and keys per thread */
  int32_t morton_ppt;
  int32_t morton_grid;
  int32_t scan_ipt;
  int32_t sort_bits;    /* digit width of a build's sort: 8 (four passes) or
                           10 (three passes over the 30-bit keys; builds with
                           the permutation keep 8) */
  int32_t sort10_kpt;   /* keys per thread of a 10-bit pass (8 or 16) */
  int32_t oct_order;    /* octree work order in a tile: 0 chains by index,
                           1 by the size of their group walk (faster on
                           clustered clouds) */
  int32_t sort_msd;     /* with sort_bits 10: 0 three LSD passes; 1 one
                           scatter by the top 10-bit digit, then each bucket
                           sorted in LDS by one workgroup; a build with a
                           bucket above 16384 keys runs the three LSD passes
                           after that, one of more than 2048 clouds only
                           them */
} ocp_tune;
int ocp_get_tune(ocp_ctx ctx, ocp_tune* out);
int ocp_set_tune(ocp_ctx ctx, const ocp_tune* in);
/* Time the supported configurations on the given workload (plain when
 * n_seg == 0, else batched) by two rounds of coordinate descent, each
 * configuration's time the best of three runs of `reps` builds, and keep
 * the fastest; writes the choice. */
int ocp_autotune(ocp_ctx ctx, const float* points, int float3, int64_t n,
                 const int32_t* seg_offsets, const float* seg_min,
                 const float* seg_range, int32_t n_seg, int reps,
                 ocp_tune* chosen);
/* ocp_autotune with its result kept in a text file, one line per workload
 * class (device architecture and compute units, plain or batched, point
 * layout, point and segment counts rounded up to powers of two): a class
 * already in the file is applied without timing anything (*from_cache = 1),
 * otherwise it is tuned and its line appended. */
int ocp_autotune_cached(ocp_ctx ctx, const char* cache_path, const float* points,
                        int float3, int64_t n, const int32_t* seg_offsets,
                        const float* seg_min, const float* seg_range,
                        int32_t n_seg, int reps, ocp_tune* chosen,
                        int* from_cache);

/* ------------------------------------------------------------ learned sort
 * No counterpart in the reference, which sorts with OneSweep only
 * (apps/gpu_demo/main.cu:63-125). A drop-in for ocp_k_SortKeys when the key
 * distribution is known from earlier frames: a small FP32 CDF model predicts
 * 16384 bucket boundaries, keys are scattered into buckets and each bucket is
 * sorted in LDS; oversized buckets are sorted by OneSweep and every bucket
 * boundary is checked, so the output bytes equal ocp_k_SortKeys' whatever
 * the model is.
 *
 * Model: root r(x) = b2 + sum_j w2[j] relu(w1[j] x + b1[j]) (16 units),
 * x = (key - kmin) / (kmax - kmin + 1), leaf l = floor(4096 r) clamped,
 * y = clamp(a[l] + b[l] (key - K_l), lo[l], lo[l+1]) with K_l the first key
 * of leaf l, bucket = floor(16384 y) clamped. Parameter layout (FP32):
 * w1[16] b1[16] w2[16] b2 | lo[4097] | a[4096] | b[4096]. */
#define OCP_CDF_PARAMS 12338
#define OCP_CDF_ROOT_PARAMS 49
#define OCP_CDF_BUCKETS 16384
#define OCP_CDF_SAMPLE_CAP (1 << 22)

typedef struct ocp_learned_stats {
  int64_t n;                  /* keys sorted */
  int32_t n_buckets_nonempty;
  int32_t max_bucket;         /* keys in the largest bucket */
  int32_t n_big;              /* buckets above 16384 keys (OneSweep fallback) */
  int32_t misordered;         /* bucket boundaries found out of order */
  int64_t keys_big;           /* keys in the fallback buckets */
  int32_t full_resort;        /* 1 if the whole array was re-sorted by OneSweep */
  int32_t fallback_passes;    /* digit passes the fallback ran: 2 over the oversized
                                 buckets whose key range varies in the low two
                                 digits only, plus 4 over the others (0, 2, 4, 6) */
} ocp_learned_stats;

/* The linear CDF over [kmin, kmax] (identity root, uniform leaves). */
int ocp_cdf_reset(ocp_ctx ctx, uint32_t kmin, uint32_t kmax);
/* Append keys[0], keys[stride], ... (device pointer) to the sample ring of
 * OCP_CDF_SAMPLE_CAP keys; the oldest samples are overwritten. */
int ocp_cdf_add_samples(ocp_ctx ctx, const uint32_t* keys, int64_t n, int32_t stride);
int ocp_cdf_clear_samples(ocp_ctx ctx);
int ocp_cdf_sample_count(ocp_ctx ctx, int64_t* count);
/* Sort the samples, train the root from the identity by `steps` SGD steps
 * (batch 1024, momentum, deterministic for a seed), fit the leaves by least
 * squares and compile the bucket boundaries. Writes the last batch loss. */
int ocp_cdf_train(ocp_ctx ctx, int32_t steps, float lr, float momentum, uint32_t seed,
                  float* loss_out);
/* Host copies of the parameters (OCP_CDF_PARAMS floats) and key range. */
int ocp_cdf_get(ocp_ctx ctx, float* params, uint32_t* kmin, uint32_t* kmax);
int ocp_cdf_set(ocp_ctx ctx, const float* params, uint32_t kmin, uint32_t kmax);
/* 32-byte header ("OCPCDF1\0", u32 version 1, u32 OCP_CDF_PARAMS, u32 kmin,
 * u32 kmax, u32 16, u32 4096) + the parameters as little-endian FP32. */
int ocp_cdf_save(ocp_ctx ctx, const char* path);
int ocp_cdf_load(ocp_ctx ctx, const char* path);
/* Mean gradient of 0.5 (r(x) - t)^2 over n device pairs (xs, ts) with
 * respect to the 49 root parameters, into a host array. */
int ocp_cdf_grad(ocp_ctx ctx, const float* xs, const float* ts, int32_t n, float* grad);
/* Bucket of each key (device arrays): mode 0 evaluates the model, mode 1
 * searches the compiled boundaries, mode 2 goes through the cell table as
 * the sort does (all three agree). */
int ocp_cdf_buckets(ocp_ctx ctx, const uint32_t* keys, uint32_t n, int mode, uint32_t* q);
/* Host copy of the 16384 boundaries: bucket(key) = #{f >= 1 : fb[f] <= key}. */
int ocp_cdf_boundaries(ocp_ctx ctx, uint32_t* fb);
/* Sort n keys in place (keys_alt: n words of scratch). Synchronous. */
int ocp_sort_learned(ocp_ctx ctx, uint32_t* keys, uint32_t* keys_alt, uint32_t n,
                     ocp_learned_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* OCTREE_PIPELINE_H */

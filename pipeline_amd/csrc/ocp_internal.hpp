// Host-side launchers of the stage kernels (one translation unit per stage)
// and the shared control block. Every launcher is asynchronous on `stream`
// and takes device pointers only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ocp_device.hpp"

namespace ocp {

// Control words in device memory (one block per context).
enum CtrlSlot : int {
  // words 0-7 are unused (tiles take their index from the workgroup id)
  kCtrlEpoch = 8,      // bumped once per launch group; tags lookback flags
  kCtrlOctTotal = 9,   // total octree slots of the batch
  kCtrlOctOk = 10,     // 1 when the octree fits the node capacity
  kCtrlMsdRun = 11,    // MSD sort: 0, or 0x7fffffff when the LSD passes must run
  kCtrlLearnBad = 12,  // learned sort: misordered bucket boundaries
  kCtrlTimeout = 13,   // mark of the launch whose wait gave up (0: none)
  kCtrlWaitLimit = 14, // 0: waits poll up to kSpinLimit times; v > 0: v - 1
  kCtrlWords = 64,
};

// Polls a bounded wait may spend before it gives up (set from the host with
// ocp_set_wait_limit; 0 makes every wait give up at once).
__device__ __forceinline__ uint32_t wait_limit(const uint32_t* ctrl) {
  const uint32_t v = ctrl[kCtrlWaitLimit];
  return v ? v - 1u : kSpinLimit;
}

// A wait that gives up stores its launch's mark (below) in kCtrlTimeout;
// the first one wins, later ones store the same mark. The keys such a pass
// leaves may no longer match the digit counts the next pass is sized by, so
// every later kernel that consumes lookback output does nothing until the
// host has reported the error and cleared the word. A kernel stops only on
// the mark of an earlier launch: that value cannot change while it runs
// (tiles of the running launch can only store its own mark), so every
// thread of every workgroup takes the same branch.
__device__ __forceinline__ uint32_t timeout_mark(uint32_t tag) { return tag | 0x80000000u; }
__device__ __forceinline__ bool stopped(const uint32_t* ctrl, uint32_t tag) {
  const uint32_t v = ctrl[kCtrlTimeout];
  return v != 0u && v != timeout_mark(tag);
}
__device__ __forceinline__ void raise_timeout(uint32_t* ctrl, uint32_t tag) {
  ctrl[kCtrlTimeout] = timeout_mark(tag);
}
// Kernels without waits stop on any mark.
__device__ __forceinline__ bool stopped_any(const uint32_t* ctrl) {
  return ctrl[kCtrlTimeout] != 0u;
}

// Lookback stage ids (part of every flag tag).
enum Stage : uint32_t {
  kStagePass0 = 0,  // OneSweep passes 0..3
  kStageUnique = 4,
  kStageScan = 5,
  kStageBucket = 6,  // learned sort: 6 and 7
};

// Kernel configuration chosen by the autotuner (or the defaults).
struct TuneTable {
  int sort_waves = 8;      // OneSweep workgroup = sort_waves * 64 threads
  int sort_kpt = 16;       // keys per thread in OneSweep
  int morton_ppt = 16;     // points per thread in the Morton kernel
  int morton_grid = 1024;  // persistent Morton workgroups
  int scan_ipt = 16;       // items per thread in unique / edge-scan tiles
  int sort_bits = 8;       // digit width of a build's sort: 8 (4 passes) or 10 (3)
  int sort10_kpt = 8;      // keys per thread of a 10-bit pass (16 waves)
  int oct_order = 0;       // octree chains in index order (0) or by group size (1)
  int sort_msd = 0;        // 10-bit digits: LSD passes (0) or MSD scatter + LDS buckets (1)
};

inline int sort_tile(const TuneTable& t) { return t.sort_waves * 64 * t.sort_kpt; }
// Digit width of a build's sort passes: 10-bit digits (three passes over
// the 30-bit Morton keys) when chosen and no permutation rides along.
inline int pipe_sort_bits(const TuneTable& t, bool perm) {
  return (t.sort_bits == 10 && !perm) ? 10 : 8;
}
// The MSD form: chosen, 10-bit digits, and few enough segments for one
// workgroup per (segment, bucket).
constexpr int32_t kMsdMaxSegments = 2048;  // grid: n_seg * 1024 workgroups of 1024
inline bool pipe_sort_msd(const TuneTable& t, bool perm, int32_t n_seg) {
  return t.sort_msd != 0 && pipe_sort_bits(t, perm) == 10 && n_seg <= kMsdMaxSegments;
}
inline int pipe_sort_tile(const TuneTable& t, bool perm) {
  return pipe_sort_bits(t, perm) == 10 ? 1024 * t.sort10_kpt : sort_tile(t);
}
// Digit-count words per segment: four 256-bin histograms for 8-bit digits,
// three 1024-bin ones for 10-bit digits.
constexpr int hist_stride(int bits) { return bits == 10 ? 3 * 1024 : 4 * 256; }
inline int morton_tile(const TuneTable& t) { return 256 * t.morton_ppt; }
inline int scan_tile(const TuneTable& t) { return 256 * t.scan_ipt; }
constexpr int kNodeTile = 256;  // BRT nodes per radix-tree / octree tile

bool sort_config_supported(int waves, int kpt);
bool sort10_config_supported(int kpt);
bool scan_config_supported(int ipt);
bool morton_config_supported(int ppt);

// ---- utility kernels (k_util.hip)
// bytes must be a multiple of 16 and both pointers 16-byte aligned.
void launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t s);
struct Affine {
  float m[12];  // rows of A (3x3), then t
};
void launch_transform(const float* src, float* dst, int64_t n, bool float3, const Affine& m,
                      float lo, float hi, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t value, int64_t n_words,
                     hipStream_t s);
// Bump the epoch and clear the learned sort's misorder count; optionally
// zero n_hist words of histogram and n2 words at z2.
void launch_begin(uint32_t* ctrl, uint32_t* hist, int64_t n_hist,
                  hipStream_t s, uint32_t* z2 = nullptr, int64_t n2 = 0);
// Begin for lookback flags shared by every context on the device: the new
// epoch is past both the context's and *dev_epoch, and is stored in both.
void launch_begin_shared(uint32_t* ctrl, uint32_t* dev_epoch, uint32_t* hist, int64_t n_hist,
                         hipStream_t s);
void launch_oct_offsets(const int32_t* kalloc, int32_t n_seg,
                        int32_t* oct_start, uint32_t* ctrl, int64_t k_cap,
                        hipStream_t s);
void launch_zero_nodes(OctNode* nodes, const uint32_t* ctrl, int64_t k_cap,
                       hipStream_t s);

// ---- input side (k_init.hip)
// n float4 points from the reference's hybrid Tausworthe generator.
void launch_init_random(float* out, int32_t n, float mn, float rg, int32_t seed, hipStream_t s);
// Per-segment cube boxes into mn[n_seg], rg[n_seg]; mm: 2 * n_seg words of
// scratch. Tiles of `tile` points per the tile map.
void launch_boxes(const float* pts, bool float3, const SegView& sv, const int32_t* tile_start,
                  int32_t total_tiles, int tile, uint32_t* mm, float* mn, float* rg,
                  hipStream_t s);

// ---- stage 1 (k_morton.hip)
// hist (optional): per-segment digit counts of the keys for hist_bits-wide
// digits (8: four 256-bin histograms; 10: three 1024-bin ones; kHistTop10:
// only the third of those, the top 10-bit digit, in the same layout).
constexpr int kHistTop10 = 1;
void launch_morton(const float* pts, bool float3, uint32_t* keys,
                   const SegView& sv, const int32_t* tile_start,
                   int32_t total_tiles, uint32_t* hist, const TuneTable& tt,
                   hipStream_t s, int hist_bits = 8);

// ---- stage 2 (k_onesweep.hip)
void launch_histogram(const uint32_t* keys, const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles,
                      uint32_t* hist, const TuneTable& tt, hipStream_t s);
// Same, over tiles of `tile` keys (a tile map built for another tile size).
void launch_histogram_tiles(const uint32_t* keys, const SegView& sv,
                            const int32_t* tile_start, int32_t total_tiles,
                            int tile, uint32_t* hist, const TuneTable& tt,
                            hipStream_t s, uint32_t* cursor0 = nullptr);
// (cursor0, optional: n_seg * 256 words zeroed for launch_cursor_pass0)
// stable = false: equal digits may leave in any order (the first pass of a
// sort of bare keys); full tiles then rank with LDS atomics. bits = 10: the
// pass takes 10-bit digits (16-wave tiles of 1024 * sort10_kpt keys, 1024
// flags per tile, hist laid out for 10-bit digits).
void launch_onesweep_pass(const uint32_t* in, uint32_t* out,
                          const SegView& sv, const int32_t* tile_start,
                          int32_t total_tiles, const uint32_t* hist,
                          uint64_t* flags, uint32_t* ctrl, int pass, bool stable,
                          const TuneTable& tt, hipStream_t s, int bits = 8,
                          const int32_t* tiles_dev = nullptr);  // (10-bit passes)
// MSD sort of bare 30-bit keys (k_msd.hip): the scatter by the top 10-bit
// digit (histogram slot 2 of the 10-bit histograms), then one workgroup per
// (segment, bucket) sorting it in LDS in place; a segment with a bucket too
// large for LDS raises ctrl[kCtrlMsdRun] instead, which gates the histogram
// of the two lower digits and the LSD passes that then sort everything
// again. Grid: n_seg * 1024 workgroups; max_seg (the largest segment's
// length) picks the LDS capacity.
void launch_msd_buckets(uint32_t* keys, const SegView& sv, const uint32_t* hist, uint32_t* ctrl,
                        int64_t max_seg, hipStream_t s);
// The scatter without a lookback: cursor holds n_seg * 1024 zeroed words
// (one per bucket), claimed by atomic adds; tiles of 1024 * sort10_kpt keys.
void launch_msd_scatter(const uint32_t* in, uint32_t* out, const SegView& sv,
                        const int32_t* tile_start, int32_t total_tiles, const uint32_t* hist,
                        uint32_t* cursor, const uint32_t* ctrl, const TuneTable& tt,
                        hipStream_t s);
// Cursor words per segment after the segments' histograms (the first sort
// pass's or the MSD scatter's: one per digit, at most 1024).
constexpr int kCursorWords = 1024;
// Digit counts of the two lower 10-bit digits (histogram slots 0 and 1),
// only when ctrl[kCtrlMsdRun] is set; tiles as the Morton kernel's.
void launch_msd_low_hist(const uint32_t* keys, const SegView& sv, const int32_t* tile_start,
                         int32_t total_tiles, uint32_t* hist, const uint32_t* ctrl,
                         const TuneTable& tt, hipStream_t s);

// First pass of an LSD sort of bare keys (digit key & 255, or key & 1023
// with bits == 10; any order among equal digits) as a cursor scatter
// (cursor_scatter.hpp): no lookback; one zeroed cursor per segment and
// digit (n_seg * 2^bits words); hist as the OneSweep passes read it (slot
// 0); tiles as OneSweep's of the tile table.
void launch_cursor_pass0(const uint32_t* in, uint32_t* out, const SegView& sv,
                         const int32_t* tile_start, const int32_t* tiles_dev, int32_t grid,
                         const uint32_t* hist, uint32_t* cursor, const uint32_t* ctrl,
                         const TuneTable& tt, hipStream_t s, int bits = 8);

// Stable pass carrying a value with every key (vin == nullptr: the key's
// index in its segment).
void launch_onesweep_pass_kv(const uint32_t* in, uint32_t* out, const uint32_t* vin,
                             uint32_t* vout, const SegView& sv, const int32_t* tile_start,
                             int32_t total_tiles, const uint32_t* hist, uint64_t* flags,
                             uint32_t* ctrl, int pass, const TuneTable& tt, hipStream_t s);

// ---- stage 3 (k_unique.hip)
// vin/vout (optional): each kept key's value (the first of its run) is
// compacted beside it.
void launch_unique(const uint32_t* in, uint32_t* out, const SegView& sv,
                   const int32_t* tile_start, int32_t total_tiles,
                   uint64_t* flags, uint32_t* ctrl, int32_t* n_unique,
                   int32_t* kalloc, const TuneTable& tt, hipStream_t s,
                   const uint32_t* vin = nullptr, uint32_t* vout = nullptr);

// ---- stage 4 (k_radix_tree.hip)
// n_brt: per-segment node counts; when n_unique != nullptr the count of
// segment k is n_unique[k]-1, else n_brt_host (single segment).
// parent_pn (optional): prefix_n[parent[i]] at i, written with parent[].
// ctrl (optional): the kernel does nothing once a lookback wait gave up.
void launch_radix_tree(const uint32_t* codes, uint8_t* prefix_n,
                       uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                       int32_t* left_child, int32_t* parent, uint8_t* parent_pn,
                       const SegView& sv,
                       const int32_t* tile_start, int32_t total_tiles,
                       const int32_t* n_unique, int32_t n_brt_host,
                       const uint32_t* ctrl, hipStream_t s);

// ---- stages 5 + 6 (k_edge_scan.hip)
// mode 0: edge count + inclusive scan fused; 1: edge count only;
// 2: inclusive scan of `edge` only.
// parent_pn (optional, fused build): used instead of prefix_n[parent[i]].
void launch_edge_scan(const uint8_t* prefix_n, const int32_t* parent,
                      const uint8_t* parent_pn, int32_t* edge, int32_t* offsets,
                      const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles,
                      const int32_t* n_unique, int32_t n_brt_host,
                      uint64_t* flags, uint32_t* ctrl, int32_t* kalloc,
                      int mode, const TuneTable& tt, hipStream_t s);

// ---- stage 7 (k_octree.hip)
struct OctArgs {
  OctNode* nodes;            // base of all segments' nodes
  const int32_t* oct_start;  // per-segment node base (nullptr: 0)
  const int32_t* kalloc;     // per-segment slot count (nullptr: k_host)
  int32_t k_host;
  const uint32_t* ctrl;      // kCtrlOctOk checked when non-null
  const uint32_t* codes;
  const uint8_t* prefix_n;
  const uint8_t* has_leaf_left;
  const uint8_t* has_leaf_right;
  const int32_t* left_child;
  const int32_t* parent;
  const int32_t* edge;
  const int32_t* offsets;
  const int32_t* n_unique;  // per segment (nullptr: n_brt_host + 1)
  int32_t n_brt_host;
};
void launch_oct_root(OctNode* nodes, const uint32_t* codes,
                     const uint8_t* prefix_n, int32_t n_brt, float mn,
                     float rg, hipStream_t s);
// The fused make + link + root pass: writes every slot, zero gaps included.
// Needs every OctArgs array (the fused build has them all); returns false
// (and launches nothing) when an array it needs is null.
bool launch_oct_build(const OctArgs& a, const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles, bool by_group,
                      hipStream_t s);

// ---- stage 7, per-stage entry points (k_octree_stage.hip)
// The make call of the reference signature has no children arrays
// (has_leaf_*, left_child may be null with phases/resolve bit 1 only).
// phase bits: 1 = make writes, 2 = link writes, 4 = leaf-mask AND-NOT
void launch_oct_scatter(const OctArgs& a, const SegView& sv,
                        const int32_t* tile_start, int32_t total_tiles,
                        int phases, hipStream_t s);
// resolve bits: 1 = resolve make keys, 2 = resolve link keys,
//               4 = write corners and cell sizes, 8 = write the root node
void launch_oct_finalize(const OctArgs& a, const SegView& sv,
                         const int32_t* tile_start, int32_t total_tiles,
                         int resolve, hipStream_t s);

// ---- node formats and validator (k_format.hip). Node counts come from
// oct_start[n_seg] on the device; boxes are per segment (smin/srg) or the
// scalars mn1/rg1.
constexpr int kTreeReportFields = 12;
void launch_seg_roots(const uint32_t* codes, const uint8_t* prefix_n, const int32_t* n_unique,
                      const int32_t* seg_start, int32_t n_seg, int32_t* root_level,
                      uint32_t* root_code, hipStream_t s);
// cov: one byte per slot, zeroed; rep: kTreeReportFields counters, zeroed.
void launch_validate(const OctNode* nodes, const int32_t* edge, const int32_t* offs,
                     const int32_t* n_unique, const int32_t* oct_start, const int32_t* seg_start,
                     int32_t n_seg, const float* smin, const float* srg, float mn1, float rg1,
                     const int32_t* root_level, const uint32_t* root_code, uint8_t* cov,
                     unsigned long long* rep, hipStream_t s);
// Scan blocks of the compact pass (bsum holds this + 1 words).
int64_t compact_scan_blocks(int64_t k_total);
// cnt[k_total] <- inclusive scan of nonzero children per slot; bsum[nb] =
// total; *bad += slots without an exact compact form (zeroed beforehand).
void launch_compact_count(const OctNode* nodes, const int32_t* oct_start, int32_t n_seg,
                          const float* smin, const float* srg, float mn1, float rg1,
                          const int32_t* root_level, int64_t k_total, int32_t* cnt,
                          uint32_t* bsum, unsigned long long* bad, hipStream_t s);
void launch_compact_write(const OctNode* nodes, const int32_t* oct_start, int32_t n_seg,
                          const float* smin, const float* srg, float mn1, float rg1,
                          const int32_t* root_level, const int32_t* incl, void* out,
                          int32_t* children, hipStream_t s);
void launch_expand(const void* in, const int32_t* children, int64_t n_children,
                   const int32_t* oct_start, int32_t n_seg, const float* smin, const float* srg,
                   float mn1, float rg1, OctNode* out, hipStream_t s);

}  // namespace ocp

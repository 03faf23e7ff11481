// Learned-CDF bucket sort: model layout, device tables and launchers
// (k_learned.hip). All launchers are asynchronous on `s`.
#pragma once

#include "ocp_internal.hpp"

namespace ocp {

// ---- model: FP32 parameters, flat
constexpr int kCdfHidden = 16;
constexpr int kCdfW1 = 0;            // [16] first-layer weights
constexpr int kCdfB1 = 16;           // [16] first-layer biases
constexpr int kCdfW2 = 32;           // [16] second-layer weights
constexpr int kCdfB2 = 48;           // [1]  output bias
constexpr int kCdfRootParams = 49;
constexpr int kCdfLeaves = 4096;
constexpr int kCdfLeafLo = 49;                         // [4097] CDF at each leaf start
constexpr int kCdfLeafA = kCdfLeafLo + kCdfLeaves + 1; // [4096] leaf intercepts
constexpr int kCdfLeafB = kCdfLeafA + kCdfLeaves;      // [4096] leaf slopes (per key)
constexpr int kCdfParams = kCdfLeafB + kCdfLeaves;     // 12338
// meta words on device: [0] kmin, [1] kmax (the key range the root is scaled to)

// ---- buckets
constexpr int kLearnBuckets = 16384;  // q in [0, 16384)
constexpr int kLearnCoarse = 256;     // first scatter digit  q >> 6
constexpr int kLearnFine = 64;        // second scatter digit q & 63
constexpr int kLearnSmallCap = 4096;  // bucket-sort classes (keys per bucket)
constexpr int kLearnMidCap = 8192;
constexpr int kLearnLargeCap = 12288;
constexpr int kLearnBigCap = 16384;   // larger buckets go to the OneSweep fallback
constexpr int kLearnPassWaves = 8, kLearnPassKpt = 16;
constexpr int kLearnPassTile = kLearnPassWaves * 64 * kLearnPassKpt;
constexpr int kLearnHistGrid = 256;   // histogram workgroups (rows)
constexpr int kLearnReduceGroups = 8; // row groups of the column sums
// Cell table: the key range [kmin, kmin + 8192 << cshift) in 8192 equal
// cells; ctab[c] = bucket of the cell's first key, so a key's bucket lies in
// [ctab[c], ctab[c+1]] and most keys need no search at all.
constexpr int kLearnCells = 8192;
constexpr int kLearnCellBits = 13;  // log2(kLearnCells)
// meta words on device: [0] kmin, [1] kmax, [2] cshift

struct LearnTables {
  const uint32_t* fb;   // [16384] bucket boundaries: q(key) >= f  <=>  key >= fb[f]
  const uint16_t* ctab; // [4097] bucket of each cell's first key
  const uint32_t* meta; // kmin, kmax, cshift
  uint32_t* rows;       // [kLearnHistGrid][16384] per-workgroup bucket counts
  uint32_t* hist;       // [kLearnReduceGroups][16384] partial sums; row 0 ends as the total
  int32_t* bstart;      // [16385] bucket starts
  uint32_t* hist_a;     // [1024] first scatter's digit counts (256 used)
  uint32_t* hist_b;     // [256 * 1024] second scatter's counts per coarse segment (64 used)
  int32_t* seg_start;   // [257] coarse segment starts
  int32_t* tile_start;  // [257] second scatter's tile map
  int32_t* tiles_b;     // [1]
  int32_t* big_start;   // [big_cap] oversized buckets, as segments
  int32_t* big_end;     // [big_cap]
  int32_t* big_tiles;   // [big_cap + 2] their tile maps, group A's then group B's
  int32_t* info;        // [8] see LearnInfo
  uint32_t* big_hist;   // [big_cap * 1024] digit histograms of the oversized buckets
  uint32_t* cursor;     // [kLearnCursorWords] the scatters' claim counters (zeroed per sort)
  uint32_t* big_cursor; // [big_cap * 256] the fallback's first-pass claim counters
  int32_t big_cap;
};
// First scatter: one cursor per coarse digit; second: 256 per coarse segment
// (the fine digit is an 8-bit digit of which 64 values occur).
constexpr int kLearnCursorWords = kLearnCoarse + kLearnCoarse * 256;
enum LearnInfo : int {
  kInfoBig = 0,         // oversized buckets
  kInfoBigTiles = 1,    // their fallback tiles
  kInfoBigKeys = 2,     // keys in them
  kInfoMaxBucket = 3,   // largest bucket
  kInfoNonEmpty = 4,    // non-empty buckets
  kInfoVary = 5,        // OR of fb[q] ^ (fb[q+1] - 1) over oversized buckets q
  kInfoBigA = 6,        // oversized buckets spanning two digits or fewer (group A)
  kInfoTilesA = 7,      // group A's fallback tiles
};

// ---- sort
void launch_lhist(const uint32_t* keys, int64_t n, const LearnTables& tb, int32_t* n_rows,
                  hipStream_t s);
void launch_lprefix(const LearnTables& tb, int32_t n_rows, int32_t fallback_tile,
                    hipStream_t s);
void launch_lpass_a(const uint32_t* in, uint32_t* out, int32_t n, const LearnTables& tb,
                    uint64_t* flags, uint32_t* ctrl, hipStream_t s);
void launch_lpass_b(const uint32_t* in, uint32_t* out, int32_t n, const LearnTables& tb,
                    uint64_t* flags, uint32_t* ctrl, hipStream_t s);
int32_t learn_pass_b_max_tiles(int32_t n);
void launch_lsort(uint32_t* keys, const LearnTables& tb, hipStream_t s);
void launch_lverify(const uint32_t* keys, const LearnTables& tb, int32_t n, uint32_t* ctrl,
                    hipStream_t s);

// ---- training and compilation
void launch_lsample(const uint32_t* keys, int64_t n, int32_t stride, uint32_t* ring,
                    int64_t ring_cap, int64_t ring_pos, int64_t count, hipStream_t s);
// meta[0..1] = sorted[0], sorted[ns-1]
void launch_lmeta(const uint32_t* sorted, int32_t ns, uint32_t* meta, hipStream_t s);
void launch_ltrain_root(const uint32_t* sorted, int32_t ns, const uint32_t* meta,
                        int32_t steps, float lr, float momentum, uint32_t seed,
                        float* model, float* loss, hipStream_t s);
void launch_lleaf_starts(const float* model, const uint32_t* meta, uint32_t* kleaf,
                         hipStream_t s);
void launch_lfit_leaves(const uint32_t* sorted, int32_t ns, const uint32_t* kleaf,
                        float* model, hipStream_t s);
// fb, then cshift (meta[2]) and the cell table
void launch_lboundaries(const float* model, uint32_t* meta, const uint32_t* kleaf,
                        uint32_t* fb, uint16_t* ctab, hipStream_t s);
// identity root over [meta[0], meta[1]], uniform leaves (the linear CDF)
void launch_llinear(float* model, hipStream_t s);
void launch_llinear_leaves(const uint32_t* kleaf, const uint32_t* meta, float* model,
                           hipStream_t s);

// ---- diagnostics
void launch_lgrad(const float* xs, const float* ts, int32_t n, const float* model,
                  float* grad, hipStream_t s);
// q from the model itself (mode 0), from a binary search of the compiled
// boundaries (mode 1) or through the cell table (mode 2, the sort's way)
void launch_lbucket(const uint32_t* keys, int32_t n, const float* model, const uint32_t* meta,
                    const uint32_t* kleaf, const uint32_t* fb, const uint16_t* ctab, int mode,
                    uint32_t* q, hipStream_t s);

}  // namespace ocp

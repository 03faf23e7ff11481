// Stage 2: OneSweep LSD radix sort of 32-bit keys, 8-bit digits, 4 passes
// (the role of k_GlobalHistogram_WithLogicalBlocks and
// k_DigitBinning_WithLogicalBlocks, kernels/cuda/02_sort.cu:208-792; the CPU
// oracle is kernels/02_sort.cpp:13-58 - keys-only, so any stable LSD sort
// gives the same bytes).
//
// One pass = one launch of onesweep_kernel (onesweep_impl.hpp) over tiles of
// WAVES*64*KPT keys with the radix digit (key >> 8*pass) & 255.
#include "cursor_scatter.hpp"
#include "onesweep_impl.hpp"

namespace ocp {
namespace {

// Digit histograms of existing keys (the sort-only entry point; inside the
// pipeline the Morton kernel produces them). Persistent workgroups walk a
// contiguous run of tiles; each thread first loads 16 keys (16 loads in
// flight: one outstanding load per thread leaves HBM idle), then counts
// them into LDS counters replicated 8 ways by lane, so that lanes of a wave
// adding the same digit (the top byte of 30-bit keys has only 64 values)
// rarely hit the same word.
constexpr int kHistRep = 8;
constexpr int kHistLoads = 16;

__global__ __launch_bounds__(256) void histogram_kernel(
    const uint32_t* __restrict__ keys, SegView sv,
    const int32_t* __restrict__ tile_start, int32_t total_tiles, int tile,
    uint32_t* __restrict__ hist, uint32_t* __restrict__ cursor0) {
  __shared__ uint32_t s_hist[1024 * kHistRep];
  const int tid = threadIdx.x;
  const uint32_t rep = lane_id() & (kHistRep - 1);
  const int32_t per = (total_tiles + gridDim.x - 1) / gridDim.x;
  const int32_t t0 = blockIdx.x * per;
  const int32_t t1 = min(total_tiles, t0 + per);
  for (int d = tid; d < 1024 * kHistRep; d += 256) s_hist[d] = 0;
  __syncthreads();
  int32_t cur = -1;
  auto flush = [&](int32_t seg) {
    __syncthreads();
    for (int d = tid; d < 1024; d += 256) {
      uint32_t v = 0;
#pragma unroll
      for (int r = 0; r < kHistRep; ++r) {
        v += s_hist[d * kHistRep + r];
        s_hist[d * kHistRep + r] = 0;
      }
      if (v) atomicAdd(&hist[static_cast<int64_t>(seg) * 1024 + d], v);
    }
    __syncthreads();
  };
  for (int32_t t = t0; t < t1; ++t) {
    const TileRef r = find_tile(tile_start, sv.n_seg, t);
    if (r.seg != cur && cur >= 0) flush(cur);
    cur = r.seg;
    // (cursor0: the first-pass cursors of the segment, zeroed by the
    // workgroup that counts its first tile)
    if (cursor0 && r.local == 0) cursor0[static_cast<int64_t>(r.seg) * 256 + tid] = 0u;
    const int32_t base = seg_begin(sv, r.seg) + r.local * tile;
    const int32_t cnt = min(tile, seg_end(sv, r.seg) - base);
    const uint32_t* src = keys + base;
    for (int32_t i0 = 0; i0 < cnt; i0 += 256 * kHistLoads) {
      uint32_t k[kHistLoads];
#pragma unroll
      for (int u = 0; u < kHistLoads; ++u) {
        const int32_t idx = i0 + u * 256 + tid;
        k[u] = idx < cnt ? src[idx] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kHistLoads; ++u) {
        if (i0 + u * 256 + tid < cnt) {
          atomicAdd(&s_hist[(k[u] & 255u) * kHistRep + rep], 1u);
          atomicAdd(&s_hist[(256 + ((k[u] >> 8) & 255u)) * kHistRep + rep], 1u);
          atomicAdd(&s_hist[(512 + ((k[u] >> 16) & 255u)) * kHistRep + rep], 1u);
          atomicAdd(&s_hist[(768 + (k[u] >> 24)) * kHistRep + rep], 1u);
        }
      }
    }
  }
  if (cur >= 0) flush(cur);
}

template <int WAVES, int KPT>
void launch_cfg(const uint32_t* in, uint32_t* out, const SegView& sv,
                const int32_t* tile_start, int32_t total_tiles,
                const uint32_t* hist, uint64_t* flags, uint32_t* ctrl,
                int pass, bool stable, hipStream_t s) {
  const uint32_t stage = kStagePass0 + static_cast<uint32_t>(pass);
  const RadixDigit dig{8u * static_cast<uint32_t>(pass)};
  if (stable)
    onesweep_kernel<WAVES, KPT, RadixDigit, true><<<total_tiles, WAVES * 64, 0, s>>>(
        in, out, sv, tile_start, nullptr, hist, pass, flags, ctrl, stage, dig);
  else
    onesweep_kernel<WAVES, KPT, RadixDigit, false><<<total_tiles, WAVES * 64, 0, s>>>(
        in, out, sv, tile_start, nullptr, hist, pass, flags, ctrl, stage, dig);
}

// 10-bit digits: one thread per digit, so 16 waves per tile.
// `digit` selects the bits (10 * digit) and the histogram slot; `stage` tags
// the lookback flags; tiles_dev (optional) is the device-side tile count.
template <int KPT>
void launch_cfg10(const uint32_t* in, uint32_t* out, const SegView& sv,
                  const int32_t* tile_start, int32_t total_tiles, const uint32_t* hist,
                  uint64_t* flags, uint32_t* ctrl, int digit, uint32_t stage, bool stable,
                  const int32_t* tiles_dev, hipStream_t s) {
  const RadixDigit10 dig{10u * static_cast<uint32_t>(digit)};
  if (stable)
    onesweep_kernel<16, KPT, RadixDigit10, true><<<total_tiles, 1024, 0, s>>>(
        in, out, sv, tile_start, tiles_dev, hist, digit, flags, ctrl, stage, dig);
  else
    onesweep_kernel<16, KPT, RadixDigit10, false><<<total_tiles, 1024, 0, s>>>(
        in, out, sv, tile_start, tiles_dev, hist, digit, flags, ctrl, stage, dig);
}

template <int WAVES, int KPT>
void launch_cfg_kv(const uint32_t* in, uint32_t* out, const uint32_t* vin, uint32_t* vout,
                   const SegView& sv, const int32_t* tile_start, int32_t total_tiles,
                   const uint32_t* hist, uint64_t* flags, uint32_t* ctrl, int pass,
                   hipStream_t s) {
  const uint32_t stage = kStagePass0 + static_cast<uint32_t>(pass);
  const RadixDigit dig{8u * static_cast<uint32_t>(pass)};
  onesweep_kernel<WAVES, KPT, RadixDigit, true, true><<<total_tiles, WAVES * 64, 0, s>>>(
      in, out, sv, tile_start, nullptr, hist, pass, flags, ctrl, stage, dig, vin, vout);
}

}  // namespace

void launch_onesweep_pass_kv(const uint32_t* in, uint32_t* out, const uint32_t* vin,
                             uint32_t* vout, const SegView& sv, const int32_t* tile_start,
                             int32_t total_tiles, const uint32_t* hist, uint64_t* flags,
                             uint32_t* ctrl, int pass, const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
#define OCP_SORT_KV_CFG(W, K)                                                         \
  if (tt.sort_waves == W && tt.sort_kpt == K) {                                       \
    launch_cfg_kv<W, K>(in, out, vin, vout, sv, tile_start, total_tiles, hist, flags, \
                        ctrl, pass, s);                                               \
    return;                                                                           \
  }
  OCP_SORT_KV_CFG(4, 16)
  OCP_SORT_KV_CFG(4, 24)
  OCP_SORT_KV_CFG(8, 8)
  OCP_SORT_KV_CFG(8, 12)
  OCP_SORT_KV_CFG(8, 16)
  OCP_SORT_KV_CFG(16, 8)
#undef OCP_SORT_KV_CFG
  launch_cfg_kv<8, 15>(in, out, vin, vout, sv, tile_start, total_tiles, hist, flags, ctrl,
                       pass, s);
}

bool sort10_config_supported(int kpt) { return kpt == 8 || kpt == 16; }

bool sort_config_supported(int waves, int kpt) {
  return (waves == 4 && (kpt == 16 || kpt == 24)) ||
         (waves == 8 && (kpt == 8 || kpt == 12 || kpt == 15 || kpt == 16)) ||
         (waves == 16 && kpt == 8);
}

void launch_onesweep_pass(const uint32_t* in, uint32_t* out,
                          const SegView& sv, const int32_t* tile_start,
                          int32_t total_tiles, const uint32_t* hist,
                          uint64_t* flags, uint32_t* ctrl, int pass, bool stable,
                          const TuneTable& tt, hipStream_t s, int bits,
                          const int32_t* tiles_dev) {
  if (total_tiles <= 0) return;
  if (bits == 10) {
    const uint32_t stage = kStagePass0 + static_cast<uint32_t>(pass);
    if (tt.sort10_kpt == 16)
      launch_cfg10<16>(in, out, sv, tile_start, total_tiles, hist, flags, ctrl, pass, stage,
                       stable, tiles_dev, s);
    else
      launch_cfg10<8>(in, out, sv, tile_start, total_tiles, hist, flags, ctrl, pass, stage,
                      stable, tiles_dev, s);
    return;
  }
#define OCP_SORT_CFG(W, K)                                                   \
  if (tt.sort_waves == W && tt.sort_kpt == K) {                              \
    launch_cfg<W, K>(in, out, sv, tile_start, total_tiles, hist, flags, ctrl, \
                     pass, stable, s);                                       \
    return;                                                                  \
  }
  OCP_SORT_CFG(4, 16)
  OCP_SORT_CFG(4, 24)
  OCP_SORT_CFG(8, 8)
  OCP_SORT_CFG(8, 12)
  OCP_SORT_CFG(8, 16)
  OCP_SORT_CFG(16, 8)
#undef OCP_SORT_CFG
  launch_cfg<8, 15>(in, out, sv, tile_start, total_tiles, hist, flags, ctrl,
                    pass, stable, s);
}

void launch_cursor_pass0(const uint32_t* in, uint32_t* out, const SegView& sv,
                        const int32_t* tile_start, const int32_t* tiles_dev, int32_t grid,
                        const uint32_t* hist, uint32_t* cursor, const uint32_t* ctrl,
                        const TuneTable& tt, hipStream_t s, int bits) {
  if (grid <= 0) return;
  if (bits == 10) {
    const RadixDigit10 d10{0u};
    if (tt.sort10_kpt == 16)
      cursor_scatter_kernel<1024, 16, RadixDigit10><<<grid, 1024, 0, s>>>(
          in, out, sv, tile_start, tiles_dev, hist, hist_stride(10), cursor, ctrl, d10);
    else
      cursor_scatter_kernel<1024, 8, RadixDigit10><<<grid, 1024, 0, s>>>(
          in, out, sv, tile_start, tiles_dev, hist, hist_stride(10), cursor, ctrl, d10);
    return;
  }
  const RadixDigit dig{0u};
#define OCP_CURSOR_CFG(W, K)                                                              \
  if (tt.sort_waves == W && tt.sort_kpt == K) {                                           \
    cursor_scatter_kernel<W * 64, K, RadixDigit><<<grid, W * 64, 0, s>>>(                 \
        in, out, sv, tile_start, tiles_dev, hist, hist_stride(8), cursor, ctrl, dig);     \
    return;                                                                               \
  }
  OCP_CURSOR_CFG(4, 16)
  OCP_CURSOR_CFG(4, 24)
  OCP_CURSOR_CFG(8, 8)
  OCP_CURSOR_CFG(8, 12)
  OCP_CURSOR_CFG(8, 16)
  OCP_CURSOR_CFG(16, 8)
#undef OCP_CURSOR_CFG
  cursor_scatter_kernel<8 * 64, 15, RadixDigit><<<grid, 8 * 64, 0, s>>>(
      in, out, sv, tile_start, tiles_dev, hist, hist_stride(8), cursor, ctrl, dig);
}

void launch_histogram(const uint32_t* keys, const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles,
                      uint32_t* hist, const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  histogram_kernel<<<grid, 256, 0, s>>>(keys, sv, tile_start, total_tiles,
                                        morton_tile(tt), hist, nullptr);
}

void launch_histogram_tiles(const uint32_t* keys, const SegView& sv,
                            const int32_t* tile_start, int32_t total_tiles,
                            int tile, uint32_t* hist, const TuneTable& tt,
                            hipStream_t s, uint32_t* cursor0) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  histogram_kernel<<<grid, 256, 0, s>>>(keys, sv, tile_start, total_tiles, tile,
                                        hist, cursor0);
}

}  // namespace ocp

// Stage 2: OneSweep LSD radix sort of 32-bit keys, 8-bit digits, 4 passes
// (the role of k_GlobalHistogram_WithLogicalBlocks and
// k_DigitBinning_WithLogicalBlocks, kernels/cuda/02_sort.cu:208-792; the CPU
// oracle is kernels/02_sort.cpp:13-58 - keys-only, so any stable LSD sort
// gives the same bytes).
//
// One pass = one launch of onesweep_kernel (onesweep_impl.hpp) over tiles of
// WAVES*64*KPT keys with the radix digit (key >> 8*pass) & 255.
#include "onesweep_impl.hpp"

namespace ocp {
namespace {

// Digit histograms of existing keys (the sort-only entry point; inside the
// pipeline the Morton kernel produces them). Same persistent structure as
// the Morton kernel.
__global__ __launch_bounds__(256) void histogram_kernel(
    const uint32_t* __restrict__ keys, SegView sv,
    const int32_t* __restrict__ tile_start, int32_t total_tiles, int tile,
    uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_hist[1024];
  const int tid = threadIdx.x;
  const int32_t per = (total_tiles + gridDim.x - 1) / gridDim.x;
  const int32_t t0 = blockIdx.x * per;
  const int32_t t1 = min(total_tiles, t0 + per);
  for (int d = tid; d < 1024; d += 256) s_hist[d] = 0;
  __syncthreads();
  int32_t cur = -1;
  auto flush = [&](int32_t seg) {
    __syncthreads();
    for (int d = tid; d < 1024; d += 256) {
      const uint32_t v = s_hist[d];
      if (v) {
        atomicAdd(&hist[static_cast<int64_t>(seg) * 1024 + d], v);
        s_hist[d] = 0;
      }
    }
    __syncthreads();
  };
  for (int32_t t = t0; t < t1; ++t) {
    const TileRef r = find_tile(tile_start, sv.n_seg, t);
    if (r.seg != cur && cur >= 0) flush(cur);
    cur = r.seg;
    const int32_t base = seg_begin(sv, r.seg) + r.local * tile;
    const int32_t cnt = min(tile, seg_end(sv, r.seg) - base);
    for (int32_t idx = tid; idx < cnt; idx += 256) {
      const uint32_t key = keys[static_cast<int64_t>(base) + idx];
      atomicAdd(&s_hist[key & 255u], 1u);
      atomicAdd(&s_hist[256 + ((key >> 8) & 255u)], 1u);
      atomicAdd(&s_hist[512 + ((key >> 16) & 255u)], 1u);
      atomicAdd(&s_hist[768 + (key >> 24)], 1u);
    }
  }
  if (cur >= 0) flush(cur);
}

template <int WAVES, int KPT>
void launch_cfg(const uint32_t* in, uint32_t* out, const SegView& sv,
                const int32_t* tile_start, int32_t total_tiles,
                const uint32_t* hist, uint64_t* flags, uint32_t* ctrl,
                int pass, hipStream_t s) {
  onesweep_kernel<WAVES, KPT, RadixDigit><<<total_tiles, WAVES * 64, 0, s>>>(
      in, out, sv, tile_start, nullptr, hist, pass, flags, ctrl,
      kStagePass0 + static_cast<uint32_t>(pass), RadixDigit{8u * static_cast<uint32_t>(pass)});
}

}  // namespace

bool sort_config_supported(int waves, int kpt) {
  return (waves == 4 && (kpt == 16 || kpt == 24)) ||
         (waves == 8 && (kpt == 8 || kpt == 12 || kpt == 15 || kpt == 16)) ||
         (waves == 16 && kpt == 8);
}

void launch_onesweep_pass(const uint32_t* in, uint32_t* out,
                          const SegView& sv, const int32_t* tile_start,
                          int32_t total_tiles, const uint32_t* hist,
                          uint64_t* flags, uint32_t* ctrl, int pass,
                          const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
#define OCP_SORT_CFG(W, K)                                                   \
  if (tt.sort_waves == W && tt.sort_kpt == K) {                              \
    launch_cfg<W, K>(in, out, sv, tile_start, total_tiles, hist, flags, ctrl, \
                     pass, s);                                               \
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
                    pass, s);
}

void launch_histogram(const uint32_t* keys, const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles,
                      uint32_t* hist, const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  histogram_kernel<<<grid, 256, 0, s>>>(keys, sv, tile_start, total_tiles,
                                        morton_tile(tt), hist);
}

void launch_histogram_tiles(const uint32_t* keys, const SegView& sv,
                            const int32_t* tile_start, int32_t total_tiles,
                            int tile, uint32_t* hist, const TuneTable& tt,
                            hipStream_t s) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  histogram_kernel<<<grid, 256, 0, s>>>(keys, sv, tile_start, total_tiles, tile,
                                        hist);
}

}  // namespace ocp

// Stage 2: OneSweep LSD radix sort of 32-bit keys, 8-bit digits, 4 passes
// (the role of k_GlobalHistogram_WithLogicalBlocks and
// k_DigitBinning_WithLogicalBlocks, kernels/cuda/02_sort.cu:208-792; the CPU
// oracle is kernels/02_sort.cpp:13-58 - keys-only, so any stable LSD sort
// gives the same bytes).
//
// One pass = one launch of onesweep_kernel over tiles of WAVES*64*KPT keys.
//   1. Tile = workgroup id. Workgroups are dispatched in id order on each
//      XCD, so the lowest unfinished tile is always resident and the
//      lookback makes progress without a shared ticket counter (a single
//      device-scope counter serialises at one atomic per tile).
//   2. Keys are loaded wave-contiguously (lane-strided, coalesced) and
//      ranked per wave with a 64-lane multisplit: eight ballots build the
//      mask of lanes holding the same digit; popcount below the lane gives a
//      stable rank; the lowest lane bumps the wave's LDS digit counter.
//   3. The tile's 256 digit counts are published at once (aggregate, or
//      inclusive for a segment's first tile), then each of 256 threads looks
//      back along its digit's chain with agent-scope loads and publishes the
//      inclusive prefix.
//   4. Keys are scattered into LDS in digit order, then written out with
//      consecutive threads on consecutive LDS slots, so equal-digit runs go
//      to consecutive HBM addresses.
// Segments sort independently: each has its own digit histogram and its own
// lookback chain, and no tile holds keys of two segments.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

template <int WAVES, int KPT>
__global__ __launch_bounds__(WAVES * 64) void onesweep_kernel(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, SegView sv,
    const int32_t* __restrict__ tile_start, const uint32_t* __restrict__ hist,
    uint64_t* __restrict__ flags, uint32_t* __restrict__ ctrl, int pass) {
  constexpr int THREADS = WAVES * 64;
  constexpr int T = THREADS * KPT;
  static_assert(THREADS >= 256, "one thread per digit is required");
  __shared__ uint32_t s_keys[T];
  __shared__ uint32_t s_whist[WAVES * 256];
  __shared__ uint32_t s_dstart[256];
  __shared__ int32_t s_dst[256];
  __shared__ uint32_t s_tmp[WAVES];

  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  for (int i = tid; i < WAVES * 256; i += THREADS) s_whist[i] = 0;
  __syncthreads();
  const int32_t t = blockIdx.x;
  const uint32_t tag = ctrl[kCtrlEpoch] * 8u + static_cast<uint32_t>(pass);
  const TileRef r = find_tile(tile_start, sv.n_seg, t);
  const int32_t sb = seg_begin(sv, r.seg);
  const int32_t base = sb + r.local * T;
  const int32_t cnt = min(T, seg_end(sv, r.seg) - base);
  const uint32_t shift = 8u * static_cast<uint32_t>(pass);

  uint32_t keys[KPT];
  uint32_t off[KPT];
  const uint32_t* src = in + base;
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int32_t idx = static_cast<int32_t>(w) * (64 * KPT) + k * 64 + static_cast<int32_t>(lane);
    keys[k] = idx < cnt ? src[idx] : 0xffffffffu;
  }

  // Wave-level multisplit. The read of the digit counter by every lane of a
  // group happens in the same LDS instruction, before the leader's update.
  uint32_t* wh = s_whist + w * 256;
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const uint32_t d = (keys[k] >> shift) & 255u;
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t rank = __popcll(m & lt);
    const uint32_t old = wh[d];
    if (rank == 0) wh[d] = old + __popcll(m);
    off[k] = old + rank;
  }
  __syncthreads();

  // Per-digit: exclusive prefix over waves, tile count, publish.
  uint32_t tile_cnt = 0;
  if (tid < 256) {
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) {
      const uint32_t c = s_whist[ww * 256 + tid];
      s_whist[ww * 256 + tid] = run;
      run += c;
    }
    if (tid == 255) run -= static_cast<uint32_t>(T - cnt);  // padding keys
    tile_cnt = run;
    flag_store(&flags[static_cast<int64_t>(t) * 256 + tid],
               flag_pack(tag, r.local == 0 ? kFlagInc : kFlagAgg, run));
  }
  uint32_t total;
  const uint32_t dstart = block_excl_scan<THREADS>(tile_cnt, s_tmp, &total);
  const uint32_t hv =
      tid < 256 ? hist[static_cast<int64_t>(r.seg) * 1024 + pass * 256 + tid] : 0u;
  const uint32_t hex = block_excl_scan<THREADS>(hv, s_tmp, &total);
  if (tid < 256) {
    s_dstart[tid] = dstart;
    uint32_t prev = 0;
    if (r.local > 0) {
      bool late = false;
      prev = lookback_n<16>(flags + tid, 256, t, t - r.local, tag, &late);
      if (late) ctrl[kCtrlTimeout] = 1u;
      flag_store(&flags[static_cast<int64_t>(t) * 256 + tid],
                 flag_pack(tag, kFlagInc, prev + tile_cnt));
    }
    s_dst[tid] = sb + static_cast<int32_t>(hex + prev) - static_cast<int32_t>(dstart);
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const uint32_t d = (keys[k] >> shift) & 255u;
    const uint32_t pos = s_dstart[d] + wh[d] + off[k];
    s_keys[pos] = keys[k];
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int32_t p = k * THREADS + tid;
    if (p < cnt) {
      const uint32_t key = s_keys[p];
      const uint32_t d = (key >> shift) & 255u;
      out[s_dst[d] + p] = key;
    }
  }
}

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
  onesweep_kernel<WAVES, KPT><<<total_tiles, WAVES * 64, 0, s>>>(
      in, out, sv, tile_start, hist, flags, ctrl, pass);
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

}  // namespace ocp

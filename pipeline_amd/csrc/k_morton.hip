// Stage 1: Morton encode (kernels/01_morton.cpp:5-14, include/shared/
// morton.h:38-46), fused with the four 8-bit digit histograms that the
// OneSweep passes need (the role of k_GlobalHistogram, kernels/cuda/
// 02_sort.cu:208-338), so the keys are read from HBM once less.
//
// Persistent workgroups of 256 threads walk a contiguous run of tiles
// (256 * PPT points each, never straddling a segment). Points arrive as
// 16-byte float4 loads (or three 4-byte loads for the float3 variant), keys
// leave as coalesced 4-byte stores. Digit counts accumulate in an LDS
// histogram (4 x 256 bins for 8-bit digits, 3 x 1024 for 10-bit ones) that is
// flushed with global atomics (zero bins skipped) when the run changes
// segment and at its end.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

// HB: digit width of the fused histograms (0: none; 8: 4 x 256 bins; 10:
// 3 x 1024 bins, for keys below 2^30; kHistTop10: the third of those only,
// the top digit that the MSD sort scatters by).
template <int PPT, bool F3, int HB>
__global__ __launch_bounds__(256) void morton_kernel(
    const float* __restrict__ pts, uint32_t* __restrict__ keys, SegView sv,
    const int32_t* __restrict__ tile_start, int32_t total_tiles,
    uint32_t* __restrict__ hist) {
  constexpr int T = 256 * PPT;
  constexpr bool HIST = HB != 0;
  constexpr bool TOP = HB == kHistTop10;
  constexpr int HW = HB == 10 ? 3072 : 1024;   // LDS histogram words
  constexpr int HS = HB == 8 ? 1024 : 3072;    // histogram words per segment
  constexpr int HO = TOP ? 2 * 1024 : 0;       // where the LDS words go in them
  __shared__ uint32_t s_hist[HIST ? HW : 1];
  const int tid = threadIdx.x;
  const int32_t per = (total_tiles + gridDim.x - 1) / gridDim.x;
  const int32_t t0 = blockIdx.x * per;
  const int32_t t1 = min(total_tiles, t0 + per);
  if (HIST) {
    for (int d = tid; d < HW; d += 256) s_hist[d] = 0;
    __syncthreads();
  }
  int32_t cur = -1;
  auto flush = [&](int32_t seg) {
    __syncthreads();
    for (int d = tid; d < HW; d += 256) {
      const uint32_t v = s_hist[d];
      if (v) {
        atomicAdd(&hist[static_cast<int64_t>(seg) * HS + HO + d], v);
        s_hist[d] = 0;
      }
    }
    __syncthreads();
  };
  for (int32_t t = t0; t < t1; ++t) {
    const TileRef r = find_tile(tile_start, sv.n_seg, t);
    if (HIST && r.seg != cur && cur >= 0) flush(cur);
    cur = r.seg;
    const int32_t sb = seg_begin(sv, r.seg);
    const int32_t base = sb + r.local * T;
    const int32_t cnt = min(T, seg_end(sv, r.seg) - base);
    const float mn = seg_min(sv, r.seg);
    const float rg = seg_range(sv, r.seg);
    auto emit = [&](int64_t g, float x, float y, float z) {
      const uint32_t key = encode(x, y, z, mn, rg);
      keys[g] = key;
      if (HB == 8) {
        atomicAdd(&s_hist[key & 255u], 1u);
        atomicAdd(&s_hist[256 + ((key >> 8) & 255u)], 1u);
        atomicAdd(&s_hist[512 + ((key >> 16) & 255u)], 1u);
        atomicAdd(&s_hist[768 + (key >> 24)], 1u);
      } else if (HB == 10) {
        atomicAdd(&s_hist[key & 1023u], 1u);
        atomicAdd(&s_hist[1024 + ((key >> 10) & 1023u)], 1u);
        atomicAdd(&s_hist[2048 + ((key >> 20) & 1023u)], 1u);
      } else if (TOP) {
        atomicAdd(&s_hist[(key >> 20) & 1023u], 1u);
      }
    };
    if (cnt == T) {
      // Full tile: every point load is issued before the first one is used
      // (a guarded load per iteration keeps one load in flight per wave).
      float px[PPT], py[PPT], pz[PPT];
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int64_t g = static_cast<int64_t>(base) + k * 256 + tid;
        if (F3) {
          px[k] = pts[3 * g + 0];
          py[k] = pts[3 * g + 1];
          pz[k] = pts[3 * g + 2];
        } else {
          const float4 p = reinterpret_cast<const float4*>(pts)[g];
          px[k] = p.x;
          py[k] = p.y;
          pz[k] = p.z;
        }
      }
#pragma unroll
      for (int k = 0; k < PPT; ++k)
        emit(static_cast<int64_t>(base) + k * 256 + tid, px[k], py[k], pz[k]);
    } else {
#pragma unroll 4
      for (int k = 0; k < PPT; ++k) {
        const int32_t idx = k * 256 + tid;
        if (idx < cnt) {
          const int64_t g = static_cast<int64_t>(base) + idx;
          if (F3) {
            emit(g, pts[3 * g + 0], pts[3 * g + 1], pts[3 * g + 2]);
          } else {
            const float4 p = reinterpret_cast<const float4*>(pts)[g];
            emit(g, p.x, p.y, p.z);
          }
        }
      }
    }
  }
  if (HIST && cur >= 0) flush(cur);
}

template <int PPT, bool F3>
void launch_f3(const float* pts, uint32_t* keys, const SegView& sv, const int32_t* tile_start,
               int32_t total_tiles, uint32_t* hist, int hist_bits, int grid, hipStream_t s) {
  if (!hist)
    morton_kernel<PPT, F3, 0><<<grid, 256, 0, s>>>(pts, keys, sv, tile_start, total_tiles, hist);
  else if (hist_bits == 10)
    morton_kernel<PPT, F3, 10><<<grid, 256, 0, s>>>(pts, keys, sv, tile_start, total_tiles, hist);
  else if (hist_bits == kHistTop10)
    morton_kernel<PPT, F3, kHistTop10><<<grid, 256, 0, s>>>(pts, keys, sv, tile_start,
                                                            total_tiles, hist);
  else
    morton_kernel<PPT, F3, 8><<<grid, 256, 0, s>>>(pts, keys, sv, tile_start, total_tiles, hist);
}

template <int PPT>
void launch_ppt(const float* pts, bool f3, uint32_t* keys, const SegView& sv,
                const int32_t* tile_start, int32_t total_tiles, uint32_t* hist,
                int hist_bits, int grid, hipStream_t s) {
  if (f3)
    launch_f3<PPT, true>(pts, keys, sv, tile_start, total_tiles, hist, hist_bits, grid, s);
  else
    launch_f3<PPT, false>(pts, keys, sv, tile_start, total_tiles, hist, hist_bits, grid, s);
}

}  // namespace

bool morton_config_supported(int ppt) {
  return ppt == 4 || ppt == 8 || ppt == 16 || ppt == 32;
}

void launch_morton(const float* pts, bool float3, uint32_t* keys,
                   const SegView& sv, const int32_t* tile_start,
                   int32_t total_tiles, uint32_t* hist, const TuneTable& tt,
                   hipStream_t s, int hist_bits) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  const int hb = hist_bits;
  switch (tt.morton_ppt) {
    case 4: launch_ppt<4>(pts, float3, keys, sv, tile_start, total_tiles, hist, hb, grid, s); break;
    case 8: launch_ppt<8>(pts, float3, keys, sv, tile_start, total_tiles, hist, hb, grid, s); break;
    case 32: launch_ppt<32>(pts, float3, keys, sv, tile_start, total_tiles, hist, hb, grid, s); break;
    default: launch_ppt<16>(pts, float3, keys, sv, tile_start, total_tiles, hist, hb, grid, s); break;
  }
}

}  // namespace ocp

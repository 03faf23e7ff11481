// MSD form of a build's sort (bare 30-bit Morton keys): one scatter by the
// top 10-bit digit (bits 20-29; cursor_scatter.hpp, no lookback: the order
// within a bucket is free) leaves every bucket of equal top digit
// contiguous, and one workgroup per bucket sorts it in LDS on the bits below
// (bucket_sort.hpp). The result is the sorted key array, the same bytes as
// the LSD passes of kernels/02_sort.cpp:5-118 give. Two global passes instead of three or four: at 10M uniform keys a
// bucket holds ~9.8K keys, well inside one workgroup's LDS.
//
// A segment with a bucket larger than the LDS capacity is not sorted here:
// the workgroups of that segment leave it alone and raise
// ctrl[kCtrlMsdRun], which the kernels that follow read: a histogram of the
// two lower 10-bit digits (the Morton kernel counts only the top one) and
// the three 10-bit LSD passes, which take it as their tile count (0: every
// tile returns at once). Those passes sort the whole array again from the
// Morton kernel's output, which the scatter left untouched, into the same
// buffer.
#include "bucket_sort.hpp"
#include "cursor_scatter.hpp"

namespace ocp {
namespace {

constexpr int kMsdRadix = 1024;

// One workgroup per (segment, bucket): blockIdx.x = seg * 1024 + digit. The
// bucket's start and size come from an exclusive scan of the segment's
// top-digit counts (hist slot 2 of the 10-bit histograms), which every
// workgroup of the segment computes alike, so all of them take the same
// decision on the segment's largest bucket.
template <int WAVES, int KPT>
__global__ __launch_bounds__(WAVES * 64) void msd_bucket_kernel(uint32_t* __restrict__ keys,
                                                                SegView sv,
                                                                const uint32_t* __restrict__ hist,
                                                                uint32_t* __restrict__ ctrl) {
  constexpr int THREADS = WAVES * 64;
  constexpr int CAP = THREADS * KPT;
  constexpr int PER = kMsdRadix / THREADS;  // digit counts per thread
  static_assert(PER * THREADS == kMsdRadix, "the counts split evenly");
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_wmax[WAVES];
  __shared__ uint32_t s_start;
  __shared__ uint32_t s_n;
  if (stopped_any(ctrl)) return;
  const int32_t seg = static_cast<int32_t>(blockIdx.x / kMsdRadix);
  const uint32_t digit = blockIdx.x % kMsdRadix;
  const int tid = threadIdx.x;
  const uint32_t* h = hist + static_cast<int64_t>(seg) * hist_stride(10) + 2 * kMsdRadix;
  uint32_t c[PER];
  uint32_t sum = 0, mx = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) c[j] = h[tid * PER + j];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    sum += c[j];
    mx = max(mx, c[j]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    mx = max(mx, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), o, 64)));
  if (lane_id() == 0) s_wmax[wave_id()] = mx;
  uint32_t total;
  const uint32_t excl = block_excl_scan<THREADS>(sum, s_scan, &total);
  if (static_cast<uint32_t>(tid) == digit / PER) {
    uint32_t st = excl;
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (static_cast<uint32_t>(j) < digit % PER) st += c[j];
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (static_cast<uint32_t>(j) == digit % PER) n = c[j];
    s_start = st;
    s_n = n;
  }
  __syncthreads();  // (the scan's own barriers already ordered s_wmax)
  uint32_t seg_max = 0;
#pragma unroll
  for (int k = 0; k < WAVES; ++k) seg_max = max(seg_max, s_wmax[k]);
  const int32_t sb = seg_begin(sv, seg);
  const int32_t len = seg_end(sv, seg) - sb;
  // Counts that do not add up to the segment cannot come from this build's
  // Morton kernel; the LSD passes then sort the segment as well.
  if (seg_max > static_cast<uint32_t>(CAP) || total != static_cast<uint32_t>(len)) {
    if (digit == 0 && tid == 0) ctrl[kCtrlMsdRun] = 0x7fffffffu;
    return;
  }
  const int32_t n = static_cast<int32_t>(s_n);
  if (n <= 1) return;
  // Every key of the bucket is digit << 20 plus 20 bits: a 12-bit first pass
  // and one stable pass after it.
  lds_bucket_sort<WAVES, KPT, 12>(keys + sb + static_cast<int32_t>(s_start), n, digit << 20, 20);
}

// Counts of the two lower 10-bit digits, for the LSD passes that run when a
// bucket did not fit: persistent workgroups over contiguous runs of the
// Morton kernel's tiles, LDS counters flushed with global atomics at each
// change of segment.
__global__ __launch_bounds__(256) void msd_low_hist_kernel(const uint32_t* __restrict__ keys,
                                                           SegView sv,
                                                           const int32_t* __restrict__ tile_start,
                                                           int32_t total_tiles, int32_t tile,
                                                           uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ ctrl) {
  __shared__ uint32_t s_hist[2 * kMsdRadix];
  if (ctrl[kCtrlMsdRun] == 0u || stopped_any(ctrl)) return;
  const int tid = threadIdx.x;
  const int32_t per = (total_tiles + gridDim.x - 1) / gridDim.x;
  const int32_t t0 = blockIdx.x * per;
  const int32_t t1 = min(total_tiles, t0 + per);
  for (int d = tid; d < 2 * kMsdRadix; d += 256) s_hist[d] = 0;
  __syncthreads();
  int32_t cur = -1;
  auto flush = [&](int32_t seg) {
    __syncthreads();
    for (int d = tid; d < 2 * kMsdRadix; d += 256) {
      const uint32_t v = s_hist[d];
      if (v) {
        atomicAdd(&hist[static_cast<int64_t>(seg) * hist_stride(10) + d], v);
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
    for (int32_t i = tid; i < cnt; i += 256) {
      const uint32_t key = keys[base + i];
      atomicAdd(&s_hist[key & 1023u], 1u);
      atomicAdd(&s_hist[kMsdRadix + ((key >> 10) & 1023u)], 1u);
    }
  }
  if (cur >= 0) flush(cur);
}

}  // namespace

void launch_msd_buckets(uint32_t* keys, const SegView& sv, const uint32_t* hist, uint32_t* ctrl,
                        int64_t max_seg, hipStream_t s) {
  if (sv.n_seg <= 0) return;
  const uint32_t grid = static_cast<uint32_t>(int64_t(sv.n_seg) * kMsdRadix);
  // Buckets of uniform keys hold max_seg / 1024 keys on average (a few
  // percent more at most for the largest): up to 11008 on average they fit
  // 12288 keys, two 512-thread workgroups per CU by LDS; larger segments
  // take 16384-key workgroups of 1024 threads, one per CU.
  if (max_seg <= int64_t(kMsdRadix) * 11008)
    msd_bucket_kernel<8, 24><<<grid, 512, 0, s>>>(keys, sv, hist, ctrl);
  else
    msd_bucket_kernel<16, 16><<<grid, 1024, 0, s>>>(keys, sv, hist, ctrl);
}

void launch_msd_scatter(const uint32_t* in, uint32_t* out, const SegView& sv,
                        const int32_t* tile_start, int32_t total_tiles, const uint32_t* hist,
                        uint32_t* cursor, const uint32_t* ctrl, const TuneTable& tt,
                        hipStream_t s) {
  if (total_tiles <= 0) return;
  // (the top digit's counts are slot 2 of each cloud's 10-bit histograms)
  const uint32_t* top = hist + 2 * kMsdRadix;
  if (tt.sort10_kpt == 16)
    cursor_scatter_kernel<1024, 16, TopDigit10><<<total_tiles, 1024, 0, s>>>(
        in, out, sv, tile_start, nullptr, top, hist_stride(10), cursor, ctrl, TopDigit10{});
  else
    cursor_scatter_kernel<1024, 8, TopDigit10><<<total_tiles, 1024, 0, s>>>(
        in, out, sv, tile_start, nullptr, top, hist_stride(10), cursor, ctrl, TopDigit10{});
}

void launch_msd_low_hist(const uint32_t* keys, const SegView& sv, const int32_t* tile_start,
                         int32_t total_tiles, uint32_t* hist, const uint32_t* ctrl,
                         const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
  const int grid = std::max(1, std::min(total_tiles, tt.morton_grid));
  msd_low_hist_kernel<<<grid, 256, 0, s>>>(keys, sv, tile_start, total_tiles, morton_tile(tt),
                                           hist, ctrl);
}

}  // namespace ocp

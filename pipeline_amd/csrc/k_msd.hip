// MSD form of a build's sort (bare 30-bit Morton keys): one OneSweep
// scatter by the top 10-bit digit (bits 20-29, k_onesweep.hip) leaves every
// bucket of equal top digit contiguous, and one workgroup per bucket sorts
// it in LDS on the bits below (bucket_sort.hpp). The result is the sorted
// key array, the same bytes as the LSD passes of kernels/02_sort.cpp:5-118
// give. Two global passes instead of three or four: at 10M uniform keys a
// bucket holds ~9.8K keys, well inside one workgroup's LDS.
//
// A segment with a bucket larger than the LDS capacity is not sorted here:
// the workgroups of that segment leave it alone and raise
// ctrl[kCtrlMsdRun], which the three 10-bit LSD passes that follow read as
// their tile count (0: every tile returns at once). Those passes sort the
// whole array again from the Morton kernel's output, which the scatter left
// untouched, into the same buffer.
#include "bucket_sort.hpp"

namespace ocp {
namespace {

constexpr int kMsdRadix = 1024;

// One workgroup per (segment, bucket): blockIdx.x = seg * 1024 + digit.
// The bucket's start and size come from an exclusive scan of the segment's
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
  lds_bucket_sort<WAVES, KPT>(keys + sb + static_cast<int32_t>(s_start), n);
}

}  // namespace

void launch_msd_buckets(uint32_t* keys, const SegView& sv, const uint32_t* hist, uint32_t* ctrl,
                        hipStream_t s) {
  if (sv.n_seg <= 0) return;
  const int64_t grid = int64_t(sv.n_seg) * kMsdRadix;
  msd_bucket_kernel<16, 16><<<static_cast<uint32_t>(grid), 1024, 0, s>>>(keys, sv, hist, ctrl);
}

}  // namespace ocp

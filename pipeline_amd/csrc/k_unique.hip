// Stage 3: unique compaction of sorted keys (kernels/03_unique.cpp:2-16:
// keep keys[i] when it differs from keys[i-1]; the reference GPU version,
// kernels/cuda/03_unique.cu:5-24, runs on one thread).
//
// Tiles of 256*IPT keys, one per workgroup. A key is kept when it is the first
// of its segment or differs from its predecessor. Per row of 64 keys a
// ballot gives the keep mask, the lane's rank (mbcnt) and the row count; one
// wave scans the 4*IPT row counts and does the two-level lookback of the
// tile total (the tile's group of 64, then the sums of earlier groups). Kept
// keys are compacted in LDS, then written with consecutive threads on
// consecutive output slots. The segment's last tile
// writes U and seeds the octree slot count (1 when U >= 1).
//
// With values (the key->point permutation), each kept key's value - the
// first of its run, i.e. the smallest point index after a stable sort - is
// compacted beside it, in place or not.
//
// In place (out == in) is safe: a tile reads all its keys, and its
// predecessor key, before publishing its count, a tile only writes below
// its own end, and it can only reach the predecessor key of the next tile
// when nothing before it was dropped, i.e. with the key that is already
// there.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

template <int IPT, bool KV>
__global__ __launch_bounds__(256) void unique_kernel(
    const uint32_t* in, uint32_t* out, SegView sv,
    const int32_t* __restrict__ tile_start, uint64_t* __restrict__ flags,
    uint32_t* __restrict__ ctrl, int32_t* __restrict__ n_unique,
    int32_t* __restrict__ kalloc, const uint32_t* vin, uint32_t* vout) {
  constexpr int THREADS = 256;
  constexpr int NW = THREADS / 64;
  constexpr int T = THREADS * IPT;
  __shared__ uint32_t s_keys[T];
  __shared__ uint32_t s_vals[KV ? T : 1];
  __shared__ uint32_t s_rows[IPT * NW];
  __shared__ uint64_t s_mask[IPT * NW];
  __shared__ uint32_t s_excl;
  __shared__ uint32_t s_total;

  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  const int32_t t = blockIdx.x;  // in-order dispatch: see lookback_n
  const uint32_t tag = ctrl[kCtrlEpoch] * 8u + kStageUnique;
  if (stopped(ctrl, tag)) return;
  const TileRef r = find_tile(tile_start, sv.n_seg, t);
  const int32_t sb = seg_begin(sv, r.seg);
  const int32_t se = seg_end(sv, r.seg);
  const int32_t base = sb + r.local * T;
  const int32_t cnt = min(T, se - base);

  // All loads are issued before any is used (a load guarded and consumed
  // in the same iteration keeps one load in flight per wave).
  uint32_t keys[IPT];
  uint32_t vals[KV ? IPT : 1];
  const uint32_t* src = in + static_cast<int64_t>(base) + tid;
  if (cnt == T) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) keys[k] = src[k * THREADS];
  } else {
#pragma unroll
    for (int k = 0; k < IPT; ++k) keys[k] = k * THREADS + tid < cnt ? src[k * THREADS] : 0u;
  }
  if constexpr (KV) {
    const uint32_t* vsrc = vin + static_cast<int64_t>(base) + tid;
#pragma unroll
    for (int k = 0; k < IPT; ++k) vals[k] = k * THREADS + tid < cnt ? vsrc[k * THREADS] : 0u;
  }
#pragma unroll
  for (int k = 0; k < IPT; ++k) s_keys[k * THREADS + tid] = keys[k];
  const uint32_t before = (r.local > 0) ? in[static_cast<int64_t>(base) - 1] : 0u;
  __syncthreads();

#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    const int32_t idx = k * THREADS + tid;
    bool keep = false;
    if (idx < cnt) {
      if (idx == 0)
        keep = (r.local == 0) || keys[k] != before;
      else
        keep = keys[k] != s_keys[idx - 1];
    }
    const uint64_t m = __ballot(keep);
    if (lane == 0) {
      s_rows[k * NW + w] = static_cast<uint32_t>(__popcll(m));
      s_mask[k * NW + w] = m;
    }
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t total = wave_scan_lds<IPT * NW>(s_rows);
    if (lane == 0)
      flag_store(&flags[t], flag_pack(tag, kFlagAgg, total));
    bool late = false;
    // Group sums live after the grid's tile flags.
    const uint32_t prev =
        wave_lookback_groups(flags, flags + gridDim.x, t, t - r.local, total, tag,
                             wait_limit(ctrl), &late);
    if (lane == 0) {
      if (late) raise_timeout(ctrl, tag);
      s_excl = prev;
      s_total = total;
      const int32_t n_tiles = (se - sb + T - 1) / T;
      if (r.local == n_tiles - 1) {
        const int32_t u = static_cast<int32_t>(prev + total);
        n_unique[r.seg] = u;
        if (kalloc) kalloc[r.seg] = u >= 1 ? 1 : 0;
      }
    }
  }
  __syncthreads();
  const uint32_t excl = s_excl;
  const uint32_t total = s_total;
  const uint64_t lt = lanemask_lt();
  // Every LDS read of a loop is issued before its results are used (a read
  // consumed in its own iteration costs a round trip per iteration).
  uint64_t mk[IPT];
  uint32_t rb[IPT];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    mk[k] = s_mask[k * NW + w];
    rb[k] = s_rows[k * NW + w];
  }
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    if ((mk[k] >> lane) & 1ull) {
      const uint32_t pos = rb[k] + static_cast<uint32_t>(__popcll(mk[k] & lt));
      s_keys[pos] = keys[k];
      if constexpr (KV) s_vals[pos] = vals[k];
    }
  }
  __syncthreads();
  uint32_t* dst = out + sb + excl;
  uint32_t kv[IPT];
  uint32_t vv[KV ? IPT : 1];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    kv[k] = s_keys[k * THREADS + tid];
    if constexpr (KV) vv[k] = s_vals[k * THREADS + tid];
  }
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    if (static_cast<uint32_t>(k * THREADS + tid) < total) {
      dst[k * THREADS + tid] = kv[k];
      if constexpr (KV) vout[sb + excl + k * THREADS + tid] = vv[k];
    }
  }
}

template <int IPT>
void launch_ipt(const uint32_t* in, uint32_t* out, const SegView& sv,
                const int32_t* tile_start, int32_t total_tiles,
                uint64_t* flags, uint32_t* ctrl, int32_t* n_unique,
                int32_t* kalloc, hipStream_t s, const uint32_t* vin, uint32_t* vout) {
  if (vin && vout)
    unique_kernel<IPT, true><<<total_tiles, 256, 0, s>>>(in, out, sv, tile_start, flags, ctrl,
                                                         n_unique, kalloc, vin, vout);
  else
    unique_kernel<IPT, false><<<total_tiles, 256, 0, s>>>(in, out, sv, tile_start, flags, ctrl,
                                                          n_unique, kalloc, nullptr, nullptr);
}

}  // namespace

bool scan_config_supported(int ipt) { return ipt == 8 || ipt == 16 || ipt == 32; }

void launch_unique(const uint32_t* in, uint32_t* out, const SegView& sv,
                   const int32_t* tile_start, int32_t total_tiles,
                   uint64_t* flags, uint32_t* ctrl, int32_t* n_unique,
                   int32_t* kalloc, const TuneTable& tt, hipStream_t s,
                   const uint32_t* vin, uint32_t* vout) {
  if (total_tiles <= 0) return;
  switch (tt.scan_ipt) {
    case 8: launch_ipt<8>(in, out, sv, tile_start, total_tiles, flags, ctrl, n_unique, kalloc, s, vin, vout); break;
    case 32: launch_ipt<32>(in, out, sv, tile_start, total_tiles, flags, ctrl, n_unique, kalloc, s, vin, vout); break;
    default: launch_ipt<16>(in, out, sv, tile_start, total_tiles, flags, ctrl, n_unique, kalloc, s, vin, vout); break;
  }
}

}  // namespace ocp

// Stages 5 + 6: edge count (kernels/05_edge_count.cpp:3-16:
// e[0] = 0, e[i] = prefix_n[i]/3 - prefix_n[parent[i]]/3) fused with the
// inclusive scan of kernels/06_partial_sum.cpp:3-15 (the reference GPU path
// has no scan and round-trips to the host, apps/gpu_demo/main.cu:308-310).
//
// Tiles of 256*IPT nodes, one per workgroup, one lookback chain per segment.
// Each thread handles one node per row of 256 (coalesced loads of prefix_n
// and, in the fused build, the parent's prefix_n stored by the radix tree;
// else parent and one gathered byte); rows are scanned with wave64 shuffles,
// one wave scans the 4*IPT row totals and does the two-level lookback (the
// tile's group of 64, then the sums of earlier groups). The tiles' maxima of
// P[i] + e[i] reach the segment's octree slot count with one atomicMax per
// group of 64 tiles.
//
// Modes: 0 = fused, 1 = edge count only, 2 = scan of `edge` only.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

template <int IPT>
__global__ __launch_bounds__(256) void edge_scan_kernel(
    const uint8_t* __restrict__ prefix_all, const int32_t* __restrict__ par_all,
    const uint8_t* __restrict__ ppn_all, int32_t* __restrict__ edge_all, int32_t* __restrict__ off_all, SegView sv,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ n_unique,
    int32_t n_brt_host, uint64_t* __restrict__ flags,
    uint32_t* __restrict__ ctrl, int32_t* __restrict__ kalloc, int mode) {
  constexpr int THREADS = 256;
  constexpr int NW = THREADS / 64;
  constexpr int T = THREADS * IPT;
  __shared__ uint32_t s_rows[IPT * NW];
  __shared__ int32_t s_max[NW];
  __shared__ uint32_t s_excl;

  const int tid = threadIdx.x;
  const uint32_t w = wave_id();
  const uint32_t lane = lane_id();
  const bool scan = mode != 1;
  const int32_t t = blockIdx.x;  // in-order dispatch: see lookback_n
  const uint32_t tag = ctrl[kCtrlEpoch] * 8u + kStageScan;
  if (stopped(ctrl, tag)) return;
  const TileRef r = find_tile(tile_start, sv.n_seg, t);
  const int32_t n = n_unique ? n_unique[r.seg] - 1 : n_brt_host;
  const int32_t sb = seg_begin(sv, r.seg);  // loaded with n, before the test
  const int32_t base = r.local * T;
  if (base >= n) return;  // no later tile of this chain waits on it
  const int32_t cnt = min(T, n - base);
  const uint8_t* pn = prefix_all + sb;
  int32_t* edge = edge_all + sb;

  // Loads first, all in flight together; then the edge counts.
  int32_t la[IPT];
  uint32_t lb[IPT];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    la[k] = 0;
    lb[k] = 0;
  }
  // One loop per source, each a uniform choice, so the loads of a loop are
  // issued together. The parent's prefix_n is stored beside the node by the
  // radix tree in the fused build, else gathered through parent[] (never
  // for node 0, whose edge count is 0).
  const int32_t i0 = base + tid;
  if (mode == 2) {
#pragma unroll
    for (int k = 0; k < IPT; ++k)
      if (k * THREADS + tid < cnt) la[k] = edge[i0 + k * THREADS];
  } else if (ppn_all && cnt == T) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      la[k] = pn[i0 + k * THREADS];
      lb[k] = ppn_all[sb + i0 + k * THREADS];
    }
  } else if (ppn_all) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      if (k * THREADS + tid < cnt) {
        la[k] = pn[i0 + k * THREADS];
        lb[k] = ppn_all[sb + i0 + k * THREADS];
      }
    }
  } else {
    int32_t par[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      par[k] = 0;
      if (k * THREADS + tid < cnt) {
        la[k] = pn[i0 + k * THREADS];
        if (i0 + k * THREADS > 0) par[k] = par_all[sb + i0 + k * THREADS];
      }
    }
#pragma unroll
    for (int k = 0; k < IPT; ++k)
      if (k * THREADS + tid < cnt && i0 + k * THREADS > 0) lb[k] = pn[par[k]];
  }
  int32_t e[IPT];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    const int32_t i = base + k * THREADS + tid;
    int32_t v = 0;
    if (k * THREADS + tid < cnt) {
      if (mode == 2)
        v = la[k];
      else if (i > 0)
        v = static_cast<int32_t>(static_cast<uint32_t>(la[k]) / 3) - static_cast<int32_t>(lb[k] / 3);
      if (mode != 2) edge[i] = v;
    }
    e[k] = v;
  }
  if (!scan) return;

  uint32_t incl[IPT];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    incl[k] = wave_incl_scan(static_cast<uint32_t>(e[k]));
    if (lane == 63) s_rows[k * NW + w] = incl[k];
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
    }
  }
  __syncthreads();
  const uint32_t excl = s_excl;
  int32_t* off = off_all + sb;
  int32_t kmax = 1;
  uint32_t rb[IPT];  // all LDS reads before the stores
#pragma unroll
  for (int k = 0; k < IPT; ++k) rb[k] = s_rows[k * NW + w];
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    const int32_t idx = k * THREADS + tid;
    if (idx < cnt) {
      const int32_t p = static_cast<int32_t>(excl + rb[k] + incl[k]);
      off[base + idx] = p;
      kmax = max(kmax, p + e[k]);
    }
  }
  if (kalloc) {
    kmax = wave_incl_max(kmax);
    if (lane == 0) s_max[w] = kmax;
    __syncthreads();
    if (w == 0) {
      int32_t m = s_max[0];
#pragma unroll
      for (int q = 1; q < NW; ++q) m = max(m, s_max[q]);
      // Tile maxima reach K_alloc one group of 64 tiles at a time: the last
      // tile of a group (or of the segment) folds in the maxima its group's
      // earlier tiles published, so the segment's counter takes one atomic
      // per group (same-address atomics from every tile serialise: 19 us at
      // 10M). The waits go to earlier tiles only.
      uint64_t* mflags = flags + 2 * static_cast<int64_t>(gridDim.x);
      if (lane == 0) flag_store(mflags + t, flag_pack(tag, kFlagAgg, static_cast<uint32_t>(m)));
      const int32_t local = r.local;
      const int32_t g0 = local & ~63;
      if ((local & 63) == 63 || local == (n + T - 1) / T - 1) {
        bool late = false;
        const uint32_t gm = wave_sum_flags<true>(mflags + (t - local) + g0, 1, local - g0,
                                                 tag & 0x3fffffffu, wait_limit(ctrl), &late);
        if (lane == 0) {
          if (late) raise_timeout(ctrl, tag);
          atomicMax(&kalloc[r.seg], max(m, static_cast<int32_t>(gm)));
        }
      }
    }
  }
}

template <int IPT>
void launch_ipt(const uint8_t* prefix_n, const int32_t* parent, const uint8_t* parent_pn,
                int32_t* edge, int32_t* offsets, const SegView& sv,
                const int32_t* tile_start, int32_t total_tiles, const int32_t* n_unique,
                int32_t n_brt_host, uint64_t* flags, uint32_t* ctrl,
                int32_t* kalloc, int mode, hipStream_t s) {
  edge_scan_kernel<IPT><<<total_tiles, 256, 0, s>>>(
      prefix_n, parent, parent_pn, edge, offsets, sv, tile_start, n_unique, n_brt_host,
      flags, ctrl, kalloc, mode);
}

}  // namespace

void launch_edge_scan(const uint8_t* prefix_n, const int32_t* parent,
                      const uint8_t* parent_pn, int32_t* edge, int32_t* offsets,
                      const SegView& sv,
                      const int32_t* tile_start, int32_t total_tiles,
                      const int32_t* n_unique, int32_t n_brt_host,
                      uint64_t* flags, uint32_t* ctrl, int32_t* kalloc,
                      int mode, const TuneTable& tt, hipStream_t s) {
  if (total_tiles <= 0) return;
  switch (tt.scan_ipt) {
    case 32: launch_ipt<32>(prefix_n, parent, parent_pn, edge, offsets, sv, tile_start, total_tiles, n_unique, n_brt_host, flags, ctrl, kalloc, mode, s); break;
    case 8: launch_ipt<8>(prefix_n, parent, parent_pn, edge, offsets, sv, tile_start, total_tiles, n_unique, n_brt_host, flags, ctrl, kalloc, mode, s); break;
    default: launch_ipt<16>(prefix_n, parent, parent_pn, edge, offsets, sv, tile_start, total_tiles, n_unique, n_brt_host, flags, ctrl, kalloc, mode, s); break;
  }
}

}  // namespace ocp

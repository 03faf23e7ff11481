// The input side of the pipeline on the device: the reference's GPU point
// generator and the per-cloud bounding box the Morton stage quantises with.
//
// - init_random_kernel: the hybrid LCG-Tausworthe generator of
//   kernels/cuda/00_init.cu:9-50 (Howes and Thomas, GPU Gems 3, chapter 37),
//   one step per point from state seeded by the point index, which is what
//   the reference computes when it launches at least n threads
//   (apps/gpu_demo/main.cu:167-173). Coordinates are (x & 0xffffff) / 2^24
//   scaled to [min, min + range); w is drawn as well, as there.
// - box kernels: per cloud, the minimum and maximum over x, y and z, then
//   the cube box (min, range) with range = max(extent * 1.0001, 1e-3), so
//   that the far face quantises below 1024 (the reference fixes the box at
//   [0, 1024), benchmarks/config.hpp:7-10; batched clouds each have their
//   own). Floats are reduced as order-preserving unsigned integers so that
//   workgroups combine their partial results with integer atomics.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

__device__ __forceinline__ uint32_t taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m) {
  return ((z & m) << s3) ^ (((z << s1) ^ z) >> s2);
}
__device__ __forceinline__ float unit24(uint32_t x) {
  return static_cast<float>(x & 0xffffffu) / 16777216.0f;
}

__global__ __launch_bounds__(256) void init_random_kernel(float4* __restrict__ out, int32_t n,
                                                          float mn, float rg, int32_t seed) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    const uint32_t idx = static_cast<uint32_t>(i);
    const uint32_t sd = static_cast<uint32_t>(seed);
    uint32_t z1 = (idx << 2) * sd;
    uint32_t z2 = ((idx << 2) + 1u) * sd;
    uint32_t z3 = ((idx << 2) + 2u) * sd;
    uint32_t z4 = ((idx << 2) + 3u) * sd;
    z1 = taus_step(z1, 13, 19, 12, 4294967294u);
    z2 = taus_step(z2, 2, 25, 4, 4294967288u);
    z3 = taus_step(z3, 3, 11, 17, 4294967280u);
    z4 = z4 * 1664525u + 1013904223u;
    float4 p;
    p.x = mn + unit24(z1 ^ z2) * rg;
    p.y = mn + unit24(z2 ^ z3) * rg;
    p.z = mn + unit24(z3 ^ z4) * rg;
    p.w = mn + unit24(z1 ^ z4) * rg;
    out[i] = p;
  }
}

// Order-preserving map of a float to an unsigned integer (NaN aside).
__device__ __forceinline__ uint32_t ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

constexpr int kBoxThreads = 256;
constexpr int kBoxPts = 8;  // points per thread per step

// Persistent workgroups over contiguous runs of point tiles (the Morton tile
// map); a run's partial minimum and maximum are folded into its cloud's two
// words with one atomic each whenever the run moves to the next cloud.
// mm: per cloud {min, max} as ord() words, set to {~0, 0} beforehand.
__global__ __launch_bounds__(kBoxThreads) void box_reduce_kernel(
    const float* __restrict__ pts, int32_t width, SegView sv,
    const int32_t* __restrict__ tile_start, int32_t total_tiles, int32_t tile,
    uint32_t* __restrict__ mm) {
  __shared__ uint32_t s_mn[kBoxThreads / 64], s_mx[kBoxThreads / 64];
  const int32_t per = (total_tiles + gridDim.x - 1) / gridDim.x;
  const int32_t t0 = blockIdx.x * per;
  const int32_t t1 = min(total_tiles, t0 + per);
  uint32_t vmn = 0xffffffffu, vmx = 0u;
  int32_t cur = -1;
  auto flush = [&](int32_t seg) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      vmn = min(vmn, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmn), o, 64)));
      vmx = max(vmx, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmx), o, 64)));
    }
    if (lane_id() == 0) {
      s_mn[wave_id()] = vmn;
      s_mx[wave_id()] = vmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t a = s_mn[0], b = s_mx[0];
#pragma unroll
      for (int w = 1; w < kBoxThreads / 64; ++w) {
        a = min(a, s_mn[w]);
        b = max(b, s_mx[w]);
      }
      if (a <= b) {
        atomicMin(&mm[2 * seg], a);
        atomicMax(&mm[2 * seg + 1], b);
      }
    }
    __syncthreads();
    vmn = 0xffffffffu;
    vmx = 0u;
  };
  for (int32_t t = t0; t < t1; ++t) {
    const TileRef r = find_tile(tile_start, sv.n_seg, t);
    if (r.seg != cur && cur >= 0) flush(cur);
    cur = r.seg;
    const int32_t base = seg_begin(sv, r.seg) + r.local * tile;
    const int32_t cnt = min(tile, seg_end(sv, r.seg) - base);
    for (int32_t i0 = 0; i0 < cnt; i0 += kBoxThreads * kBoxPts) {
      float v[kBoxPts][3];
#pragma unroll
      for (int u = 0; u < kBoxPts; ++u) {
        const int32_t idx = i0 + u * kBoxThreads + static_cast<int32_t>(threadIdx.x);
        const int64_t pi = static_cast<int64_t>(base + min(idx, cnt - 1)) * width;
        v[u][0] = pts[pi];
        v[u][1] = pts[pi + 1];
        v[u][2] = pts[pi + 2];
      }
#pragma unroll
      for (int u = 0; u < kBoxPts; ++u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const uint32_t o = ord(v[u][c]);
          vmn = min(vmn, o);
          vmx = max(vmx, o);
        }
      }
    }
  }
  if (cur >= 0) flush(cur);
}

// mm -> (min, range) per cloud: range = max(extent * 1.0001, 1e-3) computed
// in double and rounded once, extent = max - min in float. An empty cloud
// gets (0, 1).
__global__ __launch_bounds__(256) void box_finalize_kernel(const uint32_t* __restrict__ mm,
                                                           int32_t n_seg, float* __restrict__ mn,
                                                           float* __restrict__ rg) {
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n_seg) return;
  const uint32_t a = mm[2 * k], b = mm[2 * k + 1];
  if (a > b) {
    mn[k] = 0.0f;
    rg[k] = 1.0f;
    return;
  }
  const float lo = unord(a);
  const float ext = unord(b) - lo;
  mn[k] = lo;
  rg[k] = static_cast<float>(fmax(static_cast<double>(ext) * 1.0001, 1e-3));
}

__global__ void box_init_kernel(uint32_t* __restrict__ mm, int32_t n_seg) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_seg; i += gridDim.x * blockDim.x) {
    mm[2 * i] = 0xffffffffu;
    mm[2 * i + 1] = 0u;
  }
}

}  // namespace

void launch_init_random(float* out, int32_t n, float mn, float rg, int32_t seed, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>((int64_t(n) + 255) / 256, 65536));
  init_random_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<float4*>(out), n, mn, rg, seed);
}

void launch_boxes(const float* pts, bool float3, const SegView& sv, const int32_t* tile_start,
                  int32_t total_tiles, int tile, uint32_t* mm, float* mn, float* rg,
                  hipStream_t s) {
  const int32_t n_seg = sv.n_seg;
  box_init_kernel<<<std::max(1, std::min((n_seg + 255) / 256, 1024)), 256, 0, s>>>(mm, n_seg);
  if (total_tiles > 0) {
    const int grid = std::max(1, std::min(total_tiles, 1024));
    box_reduce_kernel<<<grid, kBoxThreads, 0, s>>>(pts, float3 ? 3 : 4, sv, tile_start,
                                                   total_tiles, tile, mm);
  }
  box_finalize_kernel<<<(n_seg + 255) / 256, 256, 0, s>>>(mm, n_seg, mn, rg);
}

}  // namespace ocp

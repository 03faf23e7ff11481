// Small control kernels: fills (used instead of memset nodes so a captured
// graph holds kernels only), the per-launch-group epoch bump, the octree
// node offsets of a batch, and the node zero-fill.
#include "ocp_internal.hpp"

namespace ocp {
namespace {

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p,
                                                       uint32_t value,
                                                       int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
       i < n; i += stride)
    p[i] = value;
}

__global__ __launch_bounds__(256) void begin_kernel(uint32_t* ctrl,
                                                    uint32_t* hist,
                                                    int64_t n_hist,
                                                    uint32_t* z2, int64_t n2) {
  const int64_t tid = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (tid == 0) {
    ctrl[kCtrlEpoch] = ctrl[kCtrlEpoch] + 1;
    ctrl[kCtrlLearnBad] = 0;
    ctrl[kCtrlMsdRun] = 0;
  }
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = tid; i < n_hist; i += stride) hist[i] = 0;
  for (int64_t i = tid; i < n2; i += stride) z2[i] = 0;
}

// Begin of a sort whose lookback flags live in memory shared by every
// context on the device: the context's epoch moves past the device-wide one
// kept beside those flags, and both take the new value, so a flag left by
// another context's earlier sort never carries a tag this sort waits for.
__global__ __launch_bounds__(256) void begin_shared_kernel(uint32_t* ctrl, uint32_t* dev_epoch,
                                                           uint32_t* hist, int64_t n_hist) {
  const int64_t tid = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (tid == 0) {
    const uint32_t e = max(ctrl[kCtrlEpoch], *dev_epoch) + 1u;
    ctrl[kCtrlEpoch] = e;
    *dev_epoch = e;
  }
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = tid; i < n_hist; i += stride) hist[i] = 0;
}

// Exclusive scan of the per-segment slot counts into node offsets; one
// workgroup of 1024 threads walks the segments in chunks.
__global__ __launch_bounds__(1024) void oct_offsets_kernel(
    const int32_t* __restrict__ kalloc, int32_t n_seg,
    int32_t* __restrict__ oct_start, uint32_t* ctrl, int64_t k_cap) {
  __shared__ uint32_t tmp[16];
  int64_t carry = 0;
  for (int32_t c = 0; c < n_seg; c += 1024) {
    const int32_t k = c + static_cast<int32_t>(threadIdx.x);
    const uint32_t v = k < n_seg ? static_cast<uint32_t>(kalloc[k]) : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan<1024>(v, tmp, &total);
    if (k < n_seg) oct_start[k] = static_cast<int32_t>(carry + ex);
    carry += total;
  }
  if (threadIdx.x == 0) {
    oct_start[n_seg] = static_cast<int32_t>(carry);
    ctrl[kCtrlOctTotal] = static_cast<uint32_t>(carry);
    ctrl[kCtrlOctOk] = carry <= k_cap ? 1u : 0u;
  }
}

// Zero the first ctrl[kCtrlOctTotal] nodes with 16-byte stores. The node
// buffer is allocated with 16 bytes of tail slack, so the rounded-up range
// stays inside it.
__global__ __launch_bounds__(256) void zero_nodes_kernel(OctNode* nodes,
                                                         const uint32_t* ctrl,
                                                         int64_t k_cap) {
  if (ctrl[kCtrlOctOk] == 0) return;
  const int64_t k = ctrl[kCtrlOctTotal];
  const int64_t n16 = (k * static_cast<int64_t>(sizeof(OctNode)) + 15) / 16;
  uint4* p = reinterpret_cast<uint4*>(nodes);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
       i < n16; i += stride)
    p[i] = make_uint4(0, 0, 0, 0);
}

// Plain 16-byte-per-lane copy (device to device); also the known-size
// reference kernel that HBM traffic counters are calibrated against.
__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src,
                                                   int64_t n16) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
       i < n16; i += stride)
    dst[i] = src[i];
}

// p' = clamp(A p + t, lo, hi) for each point (A row-major 3x3, t in m[9..11]);
// the fourth float of a float4 point passes through. In place is allowed.
template <int S>
__global__ __launch_bounds__(256) void transform_kernel(const float* __restrict__ src,
                                                        float* __restrict__ dst, int64_t n,
                                                        Affine m, float lo, float hi) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += stride) {
    const float* p = src + i * S;
    const float x = p[0], y = p[1], z = p[2];
    float* q = dst + i * S;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float v = fmaf(m.m[3 * r], x, fmaf(m.m[3 * r + 1], y, fmaf(m.m[3 * r + 2], z, m.m[9 + r])));
      q[r] = fminf(fmaxf(v, lo), hi);
    }
    if (S == 4) q[3] = p[3];
  }
}

}  // namespace

void launch_transform(const float* src, float* dst, int64_t n, bool float3, const Affine& m,
                      float lo, float hi, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>((n + 255) / 256, 8192));
  if (float3)
    transform_kernel<3><<<blocks, 256, 0, s>>>(src, dst, n, m, lo, hi);
  else
    transform_kernel<4><<<blocks, 256, 0, s>>>(src, dst, n, m, lo, hi);
}

void launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  const int64_t n16 = bytes / 16;
  if (n16 <= 0) return;
  const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 8192);
  copy_kernel<<<static_cast<int>(blocks), 256, 0, s>>>(
      static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16);
}

void launch_fill_u32(uint32_t* p, uint32_t value, int64_t n_words,
                     hipStream_t s) {
  if (n_words <= 0) return;
  const int64_t blocks = std::min<int64_t>((n_words + 255) / 256, 4096);
  fill_u32_kernel<<<static_cast<int>(blocks), 256, 0, s>>>(p, value, n_words);
}

void launch_begin(uint32_t* ctrl, uint32_t* hist, int64_t n_hist,
                  hipStream_t s, uint32_t* z2, int64_t n2) {
  const int64_t blocks = std::max<int64_t>(
      1, std::min<int64_t>((std::max(n_hist, n2) + 255) / 256, 2048));
  begin_kernel<<<static_cast<int>(blocks), 256, 0, s>>>(ctrl, hist, n_hist, z2, n2);
}

void launch_begin_shared(uint32_t* ctrl, uint32_t* dev_epoch, uint32_t* hist, int64_t n_hist,
                         hipStream_t s) {
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n_hist + 255) / 256, 64));
  begin_shared_kernel<<<static_cast<int>(blocks), 256, 0, s>>>(ctrl, dev_epoch, hist, n_hist);
}

void launch_oct_offsets(const int32_t* kalloc, int32_t n_seg,
                        int32_t* oct_start, uint32_t* ctrl, int64_t k_cap,
                        hipStream_t s) {
  oct_offsets_kernel<<<1, 1024, 0, s>>>(kalloc, n_seg, oct_start, ctrl, k_cap);
}

void launch_zero_nodes(OctNode* nodes, const uint32_t* ctrl, int64_t k_cap,
                       hipStream_t s) {
  const int64_t n16 = (k_cap * static_cast<int64_t>(sizeof(OctNode)) + 15) / 16;
  const int64_t blocks =
      std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, 8192));
  zero_nodes_kernel<<<static_cast<int>(blocks), 256, 0, s>>>(nodes, ctrl,
                                                             k_cap);
}

}  // namespace ocp

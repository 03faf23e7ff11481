// C ABI of include/octree_pipeline.h: context, device-memory plumbing, the
// per-stage entry points, the fused device-resident pipeline (plain and
// batched), hipGraph capture/replay, stage timing and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/octree_pipeline.h"
#include "learned.hpp"
#include "ocp_internal.hpp"

using namespace ocp;

static_assert(sizeof(ocp_oct_node) == sizeof(OctNode), "node layout");

namespace {

constexpr int64_t kMaxPoints = (int64_t(1) << 30) - 1;  // 30-bit flag counts
constexpr int64_t kMaxSegPoints = int64_t(1) << 27;     // octree order keys

template <typename T>
struct DevBuf {
  T* p = nullptr;
  int64_t n = 0;
};

struct Workspace {
  int64_t n_cap = 0;
  int32_t seg_cap = 0;
  int64_t k_cap = 0;
  int64_t flag_cap = 0;
  uint32_t* keys_a = nullptr;
  uint32_t* keys_b = nullptr;
  uint8_t* prefix_n = nullptr;
  uint8_t* hl = nullptr;
  uint8_t* hr = nullptr;
  int32_t* lc = nullptr;
  int32_t* par = nullptr;
  uint8_t* ppn = nullptr;  // prefix_n of each node's parent (fused build)
  int32_t* edge = nullptr;
  int32_t* offs = nullptr;
  OctNode* nodes = nullptr;
  uint32_t* hist = nullptr;      // seg_cap * 1024
  uint64_t* flags = nullptr;     // flag_cap words
  int32_t* counts = nullptr;     // n_unique[seg_cap], kalloc[seg_cap]
  int32_t* oct_start = nullptr;  // seg_cap + 1
  int32_t* tables = nullptr;     // seg_start + 4 tile maps, 5 * (seg_cap+1)
  float* boxes = nullptr;        // seg_min[seg_cap], seg_range[seg_cap]
  uint32_t* box_mm = nullptr;    // per segment {min, max} while boxes are reduced
  uint32_t* vals_a = nullptr;    // key->point permutation (allocated when enabled): sorted
  uint32_t* vals_b = nullptr;    //   point indices, then each unique key's first point
  int64_t vals_cap = 0;
};

struct Plan {
  int32_t n_seg = 1;
  int64_t n_total = 0;
  int64_t max_seg = 0;        // the largest segment's length
  bool device_boxes = false;  // boxes reduced on the device into ws.boxes
  bool permutation = false;   // point indices carried through sort and unique
  SegView sv{};
  int32_t tiles_morton = 0, tiles_sort = 0, tiles_scan = 0, tiles_node = 0;
  const int32_t* ts_morton = nullptr;
  const int32_t* ts_sort = nullptr;
  const int32_t* ts_scan = nullptr;
  const int32_t* ts_node = nullptr;
};

// Kernels of one build, in launch order (per-kernel event timing).
enum KernelId : int {
  kKBegin = 0, kKMorton, kKPass0, kKPass1, kKPass2, kKPass3, kKUnique,
  kKRadixTree, kKEdgeScan, kKOctOffsets, kKOctBuild, kKernelCount
};
const char* const kKernelNames[kKernelCount] = {
    "begin", "morton_hist", "onesweep_pass0", "onesweep_pass1", "onesweep_pass2",
    "onesweep_pass3", "unique", "radix_tree", "edge_scan", "oct_offsets", "octree"};
// The sort's kernels when a build takes the MSD form (the last slot unused;
// the third is named msd_lsd_skipped when every bucket fit in LDS).
const char* const kMsdKernelNames[4] = {"msd_scatter", "msd_buckets", "msd_lsd_fallback",
                                        "unused"};

struct LastBuild {
  bool valid = false;
  bool batched = false;
  const float* pts = nullptr;
  int float3 = 0;
  int64_t n = 0;
  float mn = 0, rg = 0;
  std::vector<int32_t> offsets;
  std::vector<float> seg_min, seg_range;
};

// Learned sort: the model, its compiled tables, the key-sample ring and
// the per-sort tables (allocated on first use).
struct LearnState {
  float* model = nullptr;     // kCdfParams
  uint32_t* meta = nullptr;   // [kmin, kmax, 0, 0]
  uint32_t* kleaf = nullptr;  // kCdfLeaves + 1
  uint32_t* fb = nullptr;     // kLearnBuckets
  uint16_t* ctab = nullptr;   // kLearnCells + 2
  float* loss = nullptr;
  bool compiled = false;
  uint32_t* ring = nullptr;   // sample ring
  uint32_t* samp_a = nullptr;
  uint32_t* samp_b = nullptr;
  int64_t ring_fill = 0, ring_pos = 0;
  // per-sort tables
  int64_t n_cap = 0;
  LearnTables tb{};
  int32_t* h_info = nullptr;     // pinned host copy of tb.info
  hipEvent_t ev_info = nullptr;
};
constexpr int64_t kSampleCap = int64_t(1) << 22;

// ---- Scratch of ocp_k_SortKeys_static, the sort that allocates nothing.
// One fixed array per device, in the code object's device memory: a 256-byte
// header (word 0: the device-wide epoch of the flags below), the 4 x 256
// digit counts, then the lookback flags of up to kStaticFlagWords / 256
// tiles - enough for kStaticSortMaxKeys keys at the smallest sort tile
// (4096 keys). 8 MiB + 4.25 KiB per device, held until the process ends.
constexpr int64_t kStaticSortMaxKeys = int64_t(1) << 24;
constexpr int64_t kStaticHeaderBytes = 256;
constexpr int64_t kStaticHistBytes = 1024 * sizeof(uint32_t);
constexpr int64_t kStaticFlagWords = (kStaticSortMaxKeys / 4096) * 256;
constexpr int64_t kStaticScratchBytes =
    kStaticHeaderBytes + kStaticHistBytes + kStaticFlagWords * int64_t(sizeof(uint64_t));
static_assert((kStaticHeaderBytes + kStaticHistBytes) % 256 == 0, "flags stay 256-byte aligned");

}  // namespace

// In host code this name is only a stand-in for the symbol; the memory is
// reached through the per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current: the device address of
// `scratch`, whose flags are then cleared (no stale flag may carry a tag).
// Called from ocp_create, the library's set-up call.
int scratch_init(int device, hipStream_t stream, std::string* why) {
  if (device < 0 || device >= kMaxDevices) return *why = "device index out of range", OCP_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return OCP_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return *why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), OCP_ERR_HIP;
  launch_fill_u32(static_cast<uint32_t*>(ptr), 0u, kStaticScratchBytes / 4, stream);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return *why = std::string("scratch clear: ") + hipGetErrorString(e), OCP_ERR_HIP;
  g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
  return OCP_OK;
}

}  // namespace

struct ocp_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  TuneTable tune;
  std::string err;
  uint32_t* ctrl = nullptr;
  // per-stage scratch
  uint64_t* s_flags = nullptr;
  int64_t s_flag_cap = 0;
  uint32_t* s_hist = nullptr;
  Workspace ws;
  // batch tables currently on device
  std::vector<int32_t> tab_offsets;
  std::vector<float> tab_min, tab_rg;
  int tab_tune_key = -1;
  bool have_result = false;
  Plan last_plan;
  bool graph_mode = false;
  bool permutation = false;
  // Streaming: the first lookback launch of a build waits for order_wait,
  // and order_record is recorded after its last one (eager builds only).
  hipEvent_t order_wait = nullptr;
  hipEvent_t order_record = nullptr;
  hipError_t order_err = hipSuccess;
  // Cold timing: with stage timing on, this buffer is overwritten before
  // every timed kernel of an eager build (evicts L2 and the Infinity Cache).
  uint32_t* cold_buf = nullptr;
  int64_t cold_words = 0;
  uint32_t cold_seq = 0;
  std::map<std::string, hipGraphExec_t> graphs;
  bool stage_timing = false;
  hipEvent_t ev[8] = {};
  hipEvent_t kev[2 * kKernelCount] = {};
  bool kran[kKernelCount] = {};  // kernels the last timed eager build ran
  bool kmsd = false;             // ... with the MSD sort
  bool ev_recorded = false;
  LastBuild last;
  LearnState learn;
};

namespace {

int fail(ocp_ctx c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define OCP_HIP(c, expr)                                                   \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(c, OCP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define OCP_CHECK_LAUNCH(c) OCP_HIP(c, hipGetLastError())

template <typename T>
int dev_alloc(ocp_ctx c, T** p, int64_t n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  OCP_HIP(c, hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * static_cast<size_t>(n)));
  return OCP_OK;
}
template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

void clear_graphs(ocp_ctx c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

int ensure_scratch(ocp_ctx c, int64_t flag_words) {
  if (flag_words > c->s_flag_cap) {
    dev_free(c->s_flags);
    const int64_t cap = std::max<int64_t>(flag_words, 1 << 16);
    if (int r = dev_alloc(c, &c->s_flags, cap)) return r;
    OCP_HIP(c, hipMemsetAsync(c->s_flags, 0, sizeof(uint64_t) * cap, c->stream));
    c->s_flag_cap = cap;
  }
  return OCP_OK;
}

void free_workspace(Workspace& w) {
  dev_free(w.keys_a); dev_free(w.keys_b); dev_free(w.prefix_n); dev_free(w.hl);
  dev_free(w.hr); dev_free(w.lc); dev_free(w.par); dev_free(w.ppn); dev_free(w.edge);
  dev_free(w.offs); dev_free(w.nodes); dev_free(w.hist); dev_free(w.flags);
  dev_free(w.counts); dev_free(w.oct_start); dev_free(w.tables); dev_free(w.boxes);
  dev_free(w.box_mm); dev_free(w.vals_a); dev_free(w.vals_b);
  w = Workspace{};
}

int alloc_nodes(ocp_ctx c, int64_t k_cap) {
  Workspace& w = c->ws;
  dev_free(w.nodes);
  // 16 bytes of slack: the zero-fill kernel rounds up to 16-byte stores.
  const int64_t bytes = k_cap * static_cast<int64_t>(sizeof(OctNode)) + 16;
  OCP_HIP(c, hipMalloc(reinterpret_cast<void**>(&w.nodes), static_cast<size_t>(bytes)));
  w.k_cap = k_cap;
  return OCP_OK;
}

int ensure_workspace(ocp_ctx c, int64_t n, int32_t n_seg, int64_t flag_words) {
  Workspace& w = c->ws;
  const bool grow_n = n > w.n_cap;
  const bool grow_seg = n_seg > w.seg_cap;
  const bool grow_flags = flag_words > w.flag_cap;
  if (!grow_n && !grow_seg && !grow_flags) return OCP_OK;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  clear_graphs(c);
  c->have_result = false;
  c->tab_offsets.clear();
  if (grow_n) {
    const int64_t cap = std::max<int64_t>(n, 4096);
    dev_free(w.keys_a); dev_free(w.keys_b); dev_free(w.prefix_n); dev_free(w.hl);
    dev_free(w.hr); dev_free(w.lc); dev_free(w.par); dev_free(w.ppn); dev_free(w.edge);
    dev_free(w.offs);
    int r = 0;
    if ((r = dev_alloc(c, &w.keys_a, cap)) || (r = dev_alloc(c, &w.keys_b, cap)) ||
        (r = dev_alloc(c, &w.prefix_n, cap)) || (r = dev_alloc(c, &w.hl, cap)) ||
        (r = dev_alloc(c, &w.hr, cap)) || (r = dev_alloc(c, &w.lc, cap)) ||
        (r = dev_alloc(c, &w.par, cap)) || (r = dev_alloc(c, &w.ppn, cap)) ||
        (r = dev_alloc(c, &w.edge, cap)) ||
        (r = dev_alloc(c, &w.offs, cap)))
      return r;
    w.n_cap = cap;
    if (w.k_cap < cap)
      if (int r2 = alloc_nodes(c, cap)) return r2;
  }
  if (grow_seg) {
    const int32_t cap = std::max<int32_t>(n_seg, 1);
    dev_free(w.hist); dev_free(w.counts); dev_free(w.oct_start); dev_free(w.tables);
    dev_free(w.boxes); dev_free(w.box_mm);
    int r = 0;
    if ((r = dev_alloc(c, &w.hist, int64_t(cap) * (hist_stride(10) + kCursorWords))) ||
        (r = dev_alloc(c, &w.counts, int64_t(cap) * 2)) ||
        (r = dev_alloc(c, &w.oct_start, int64_t(cap) + 1)) ||
        (r = dev_alloc(c, &w.tables, 5 * (int64_t(cap) + 1))) ||
        (r = dev_alloc(c, &w.boxes, 2 * int64_t(cap))) ||
        (r = dev_alloc(c, &w.box_mm, 2 * int64_t(cap))))
      return r;
    w.seg_cap = cap;
  }
  if (grow_flags) {
    dev_free(w.flags);
    const int64_t cap = std::max<int64_t>(flag_words, 1 << 16);
    if (int r = dev_alloc(c, &w.flags, cap)) return r;
    OCP_HIP(c, hipMemsetAsync(w.flags, 0, sizeof(uint64_t) * cap, c->stream));
    w.flag_cap = cap;
  }
  return OCP_OK;
}

// The permutation buffers follow the workspace's point capacity.
int ensure_perm(ocp_ctx c) {
  Workspace& w = c->ws;
  if (!c->permutation || w.vals_cap >= w.n_cap) return OCP_OK;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  clear_graphs(c);
  dev_free(w.vals_a);
  dev_free(w.vals_b);
  w.vals_cap = 0;
  int r = 0;
  if ((r = dev_alloc(c, &w.vals_a, w.n_cap)) || (r = dev_alloc(c, &w.vals_b, w.n_cap))) return r;
  w.vals_cap = w.n_cap;
  return OCP_OK;
}

SegView single_view(int64_t n, float mn, float rg) {
  SegView sv{};
  sv.n_seg = 1;
  sv.start = nullptr;
  sv.mn = nullptr;
  sv.rg = nullptr;
  sv.n1 = static_cast<int32_t>(n);
  sv.min1 = mn;
  sv.range1 = rg;
  return sv;
}

int tune_key(const TuneTable& t) {
  const int k = ((t.sort_waves * 64 + t.sort_kpt) * 64 + t.morton_ppt) * 64 + t.scan_ipt;
  return (((k * 2 + (t.sort_bits == 10)) * 32 + t.sort10_kpt) * 2 + (t.oct_order != 0)) * 2 +
         (t.sort_msd != 0);
}

int64_t flag_words_for(const TuneTable& tt, int64_t n, int32_t n_seg) {
  const int64_t sort_tiles = cdiv(n, sort_tile(tt)) + n_seg;
  const int64_t sort10_tiles = cdiv(n, 1024 * tt.sort10_kpt) + n_seg;
  const int64_t scan_tiles = cdiv(n, scan_tile(tt)) + n_seg;
  // sort passes: one flag per digit and tile; scans: tile flags, group sums,
  // tile maxima
  const int64_t sort = std::max(sort_tiles * 256, tt.sort_bits == 10 ? sort10_tiles * 1024 : 0);
  return std::max(sort, 3 * scan_tiles);
}

// Plain plan: one segment, scalar box, no device tables.
Plan plain_plan(const TuneTable& tt, int64_t n, float mn, float rg, bool perm) {
  Plan p;
  p.n_seg = 1;
  p.n_total = n;
  p.max_seg = n;
  p.permutation = perm;
  p.sv = single_view(n, mn, rg);
  p.tiles_morton = static_cast<int32_t>(cdiv(n, morton_tile(tt)));
  p.tiles_sort = static_cast<int32_t>(cdiv(n, pipe_sort_tile(tt, perm)));
  p.tiles_scan = static_cast<int32_t>(cdiv(n, scan_tile(tt)));
  p.tiles_node = static_cast<int32_t>(cdiv(n, kNodeTile));
  return p;
}

// Batched plan: validates the table, uploads offsets, boxes and tile maps
// (only when they changed).
// Host-side checks of a batch table; no device state is touched.
int validate_batch(const int32_t* offsets, const float* smin, const float* srg,
                   int32_t n_seg, std::string* why) {
  if (n_seg < 1) return *why = "n_seg must be >= 1", OCP_ERR_INVALID;
  if (!offsets) return *why = "null segment table", OCP_ERR_INVALID;
  if (!smin != !srg) return *why = "seg_min and seg_range must both be given or both be NULL",
                     OCP_ERR_INVALID;
  if (offsets[0] != 0) return *why = "seg_offsets[0] must be 0", OCP_ERR_INVALID;
  for (int32_t k = 0; k < n_seg; ++k) {
    const int64_t len = int64_t(offsets[k + 1]) - offsets[k];
    if (len < 0) return *why = "seg_offsets must be non-decreasing", OCP_ERR_INVALID;
    if (len > kMaxSegPoints) return *why = "segment larger than 2^27 points", OCP_ERR_CAPACITY;
    if (srg && !(srg[k] > 0.0f)) return *why = "seg_range must be > 0", OCP_ERR_INVALID;
    if (smin && !(smin[k] == smin[k])) return *why = "seg_min must not be NaN", OCP_ERR_INVALID;
  }
  if (int64_t(offsets[n_seg]) > kMaxPoints)
    return *why = "more than 2^30-1 points", OCP_ERR_CAPACITY;
  return OCP_OK;
}

int batched_plan(ocp_ctx c, const int32_t* offsets, const float* smin,
                 const float* srg, int32_t n_seg, bool perm, Plan* out) {
  std::string why;
  if (int r = validate_batch(offsets, smin, srg, n_seg, &why)) return fail(c, r, why);
  const int64_t n = offsets[n_seg];
  const TuneTable& tt = c->tune;
  if (int r = ensure_workspace(c, n, n_seg, flag_words_for(tt, n, n_seg))) return r;
  Workspace& w = c->ws;
  const int32_t stride = w.seg_cap + 1;
  int32_t* d_start = w.tables;
  int32_t* d_tm = w.tables + stride;
  int32_t* d_ts = w.tables + 2 * stride;
  int32_t* d_tc = w.tables + 3 * stride;
  int32_t* d_tn = w.tables + 4 * stride;
  Plan p;
  p.n_seg = n_seg;
  p.n_total = n;
  p.permutation = perm;
  p.device_boxes = !smin;
  const int stile = pipe_sort_tile(tt, perm);
  std::vector<int32_t> host(5 * static_cast<size_t>(n_seg + 1));
  int64_t tm = 0, ts = 0, tc = 0, tn = 0;
  for (int32_t k = 0; k <= n_seg; ++k) {
    host[k] = offsets[k];
    host[(n_seg + 1) + k] = static_cast<int32_t>(tm);
    host[2 * (n_seg + 1) + k] = static_cast<int32_t>(ts);
    host[3 * (n_seg + 1) + k] = static_cast<int32_t>(tc);
    host[4 * (n_seg + 1) + k] = static_cast<int32_t>(tn);
    if (k < n_seg) {
      const int64_t len = int64_t(offsets[k + 1]) - offsets[k];
      p.max_seg = std::max(p.max_seg, len);
      tm += cdiv(len, morton_tile(tt));
      ts += cdiv(len, stile);
      tc += cdiv(len, scan_tile(tt));
      tn += cdiv(len, kNodeTile);
    }
  }
  const int tab_key = tune_key(tt) * 2 + (pipe_sort_bits(tt, perm) == 10);
  const bool same = c->tab_tune_key == tab_key &&
                    c->tab_offsets.size() == static_cast<size_t>(n_seg + 1) &&
                    std::equal(offsets, offsets + n_seg + 1, c->tab_offsets.begin());
  if (!same) {
    for (int q = 0; q < 5; ++q)
      OCP_HIP(c, hipMemcpyAsync(w.tables + q * stride, host.data() + q * (n_seg + 1),
                                sizeof(int32_t) * (n_seg + 1), hipMemcpyHostToDevice,
                                c->stream));
    c->tab_offsets.assign(offsets, offsets + n_seg + 1);
    c->tab_tune_key = tab_key;
    c->tab_min.clear();
  }
  const bool dev_boxes = !smin;
  const bool same_box = dev_boxes || (c->tab_min.size() == static_cast<size_t>(n_seg) &&
                                      std::equal(smin, smin + n_seg, c->tab_min.begin()) &&
                                      std::equal(srg, srg + n_seg, c->tab_rg.begin()));
  if (dev_boxes) c->tab_min.clear();  // ws.boxes is overwritten on the device
  if (!same_box) {
    OCP_HIP(c, hipMemcpyAsync(w.boxes, smin, sizeof(float) * n_seg, hipMemcpyHostToDevice, c->stream));
    OCP_HIP(c, hipMemcpyAsync(w.boxes + w.seg_cap, srg, sizeof(float) * n_seg,
                              hipMemcpyHostToDevice, c->stream));
    OCP_HIP(c, hipStreamSynchronize(c->stream));  // host arrays may go away
    c->tab_min.assign(smin, smin + n_seg);
    c->tab_rg.assign(srg, srg + n_seg);
  }
  p.sv.n_seg = n_seg;
  p.sv.start = d_start;
  p.sv.mn = w.boxes;
  p.sv.rg = w.boxes + w.seg_cap;
  p.sv.n1 = static_cast<int32_t>(n);
  p.sv.min1 = 0.0f;
  p.sv.range1 = 1.0f;
  p.tiles_morton = static_cast<int32_t>(tm);
  p.tiles_sort = static_cast<int32_t>(ts);
  p.tiles_scan = static_cast<int32_t>(tc);
  p.tiles_node = static_cast<int32_t>(tn);
  p.ts_morton = d_tm;
  p.ts_sort = d_ts;
  p.ts_scan = d_tc;
  p.ts_node = d_tn;
  *out = p;
  return OCP_OK;
}

// Brackets one kernel launch with events (eager builds with timing on).
struct KScope {
  ocp_ctx c;
  int id;
  bool on;
  KScope(ocp_ctx c_, int id_, bool capturing)
      : c(c_), id(id_), on(c_->stage_timing && !capturing) {
    if (on && c->cold_buf)
      launch_fill_u32(c->cold_buf, ++c->cold_seq, c->cold_words, c->stream);
    if (on) {
      (void)hipEventRecord(c->kev[2 * id], c->stream);
      c->kran[id] = true;
    }
  }
  ~KScope() {
    if (on) (void)hipEventRecord(c->kev[2 * id + 1], c->stream);
  }
};

OctArgs pipeline_oct_args(ocp_ctx c) {
  Workspace& w = c->ws;
  OctArgs a{};
  a.nodes = w.nodes;
  a.oct_start = w.oct_start;
  a.kalloc = w.counts + w.seg_cap;
  a.k_host = 0;
  a.ctrl = c->ctrl;
  a.codes = w.keys_b;
  a.prefix_n = w.prefix_n;
  a.has_leaf_left = w.hl;
  a.has_leaf_right = w.hr;
  a.left_child = w.lc;
  a.parent = w.par;
  a.edge = w.edge;
  a.offsets = w.offs;
  a.n_unique = w.counts;
  a.n_brt_host = 0;
  return a;
}

// False when the octree kernel refused its arguments (nothing launched).
bool enqueue_octree(ocp_ctx c, const Plan& p, bool capturing) {
  Workspace& w = c->ws;
  hipStream_t s = c->stream;
  int32_t* kalloc = w.counts + w.seg_cap;
  { KScope k(c, kKOctOffsets, capturing);
    launch_oct_offsets(kalloc, p.n_seg, w.oct_start, c->ctrl, w.k_cap, s); }
  const OctArgs a = pipeline_oct_args(c);
  KScope k(c, kKOctBuild, capturing);
  return launch_oct_build(a, p.sv, p.ts_node, p.tiles_node, c->tune.oct_order != 0, s);
}

void mark(ocp_ctx c, int k, bool capturing) {
  if (c->stage_timing && !capturing) (void)hipEventRecord(c->ev[k], c->stream);
}

// The whole chain, enqueued on the context stream. False when a launcher
// refused its arguments.
bool enqueue_pipeline(ocp_ctx c, const float* pts, bool f3, const Plan& p,
                      bool capturing) {
  Workspace& w = c->ws;
  hipStream_t s = c->stream;
  const TuneTable& tt = c->tune;
  int32_t* n_unique = w.counts;
  int32_t* kalloc = w.counts + w.seg_cap;
  // 8-bit digits take four passes over the keys, 10-bit digits three; the
  // Morton kernel writes the buffer that leaves the sorted keys in keys_a.
  const int bits = pipe_sort_bits(tt, p.permutation);
  const int passes = bits == 10 ? 3 : 4;
  const bool msd = pipe_sort_msd(tt, p.permutation, p.n_seg);
  if (!capturing) {
    for (bool& r : c->kran) r = false;
    c->kmsd = msd;
  }
  mark(c, 0, capturing);
  // Begin also clears the per-segment U and K counts (unique leaves an
  // empty segment's entries untouched).
  { KScope k(c, kKBegin, capturing);
    // (the first sort pass's, or the MSD scatter's, cursors follow the
    // clouds' histograms)
    launch_begin(c->ctrl, w.hist,
                 int64_t(p.n_seg) * (hist_stride(bits) + (p.permutation ? 0 : (1 << bits))), s,
                 reinterpret_cast<uint32_t*>(w.counts), int64_t(w.seg_cap) * 2); }
  if (p.device_boxes)
    launch_boxes(pts, f3, p.sv, p.ts_morton, p.tiles_morton, morton_tile(tt), w.box_mm, w.boxes,
                 w.boxes + w.seg_cap, s);
  uint32_t* src = (passes % 2) ? w.keys_b : w.keys_a;
  uint32_t* dst = (passes % 2) ? w.keys_a : w.keys_b;
  { KScope k(c, kKMorton, capturing);
    launch_morton(pts, f3, src, p.sv, p.ts_morton, p.tiles_morton, w.hist, tt, s,
                  msd ? kHistTop10 : bits); }
  mark(c, 1, capturing);
  uint32_t* vsrc = nullptr;  // the first pass numbers the points itself
  uint32_t* vdst = w.vals_b;
  // Lookback launches of two builds never run at the same time (each relies
  // on its lowest unfinished tile being resident); the rest may overlap.
  if (c->order_wait && !capturing && c->order_err == hipSuccess)
    c->order_err = hipStreamWaitEvent(s, c->order_wait, 0);
  if (msd) {
    // Scatter by the top digit into keys_a, sort each bucket there in LDS;
    // if a bucket was too large, the counts of the two lower digits and
    // three LSD passes from the Morton output (left intact in keys_b) sort
    // into keys_a instead.
    { KScope k(c, kKPass0, capturing);
      launch_msd_scatter(src, dst, p.sv, p.ts_sort, p.tiles_sort, w.hist,
                         w.hist + int64_t(p.n_seg) * hist_stride(10), c->ctrl, tt, s); }
    { KScope k(c, kKPass1, capturing);
      launch_msd_buckets(dst, p.sv, w.hist, c->ctrl, p.max_seg, s); }
    KScope k(c, kKPass2, capturing);
    launch_msd_low_hist(src, p.sv, p.ts_morton, p.tiles_morton, w.hist, c->ctrl, tt, s);
    const int32_t* run = reinterpret_cast<const int32_t*>(c->ctrl + kCtrlMsdRun);
    for (int pass = 0; pass < passes; ++pass) {
      launch_onesweep_pass(src, dst, p.sv, p.ts_sort, p.tiles_sort, w.hist, w.flags, c->ctrl,
                           pass, pass != 0, tt, s, bits, run);
      std::swap(src, dst);
    }
  }
  for (int pass = 0; pass < passes && !msd; ++pass) {
    KScope k(c, kKPass0 + pass, capturing);
    if (p.permutation) {
      // Point indices ride with the keys; every pass is stable, so they end
      // as the stable argsort of the keys.
      launch_onesweep_pass_kv(src, dst, vsrc, vdst, p.sv, p.ts_sort, p.tiles_sort, w.hist,
                              w.flags, c->ctrl, pass, tt, s);
      vsrc = vdst;
      vdst = (vdst == w.vals_b) ? w.vals_a : w.vals_b;
    } else if (pass == 0) {
      // The first pass may reorder equal digits (the keys carry no payload):
      // a cursor scatter, without a lookback.
      launch_cursor_pass0(src, dst, p.sv, p.ts_sort, nullptr, p.tiles_sort, w.hist,
                          w.hist + int64_t(p.n_seg) * hist_stride(bits), c->ctrl, tt, s, bits);
    } else {
      launch_onesweep_pass(src, dst, p.sv, p.ts_sort, p.tiles_sort, w.hist, w.flags,
                           c->ctrl, pass, true, tt, s, bits);
    }
    std::swap(src, dst);
  }
  mark(c, 2, capturing);
  { KScope k(c, kKUnique, capturing);
    launch_unique(w.keys_a, w.keys_b, p.sv, p.ts_scan, p.tiles_scan, w.flags, c->ctrl,
                  n_unique, kalloc, tt, s, p.permutation ? w.vals_a : nullptr,
                  p.permutation ? w.vals_b : nullptr); }
  mark(c, 3, capturing);
  { KScope k(c, kKRadixTree, capturing);
    launch_radix_tree(w.keys_b, w.prefix_n, w.hl, w.hr, w.lc, w.par, w.ppn, p.sv, p.ts_node,
                      p.tiles_node, n_unique, 0, c->ctrl, s); }
  mark(c, 4, capturing);
  { KScope k(c, kKEdgeScan, capturing);
    launch_edge_scan(w.prefix_n, w.par, w.ppn, w.edge, w.offs, p.sv, p.ts_scan, p.tiles_scan,
                     n_unique, 0, w.flags, c->ctrl, kalloc, 0, tt, s); }
  if (c->order_record && !capturing && c->order_err == hipSuccess)
    c->order_err = hipEventRecord(c->order_record, s);
  mark(c, 5, capturing);
  const bool ok = enqueue_octree(c, p, capturing);
  mark(c, 6, capturing);
  return ok;
}

std::string graph_key(const char* kind, const float* pts, int f3, int64_t n,
                      float mn, float rg, const std::vector<int32_t>* offs,
                      const TuneTable& tt) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s|%p|%d|%lld|%a|%a|%d", kind,
                static_cast<const void*>(pts), f3, static_cast<long long>(n), mn, rg,
                tune_key(tt));
  std::string k(buf);
  if (offs) {
    uint64_t h = 1469598103934665603ull;
    for (int32_t v : *offs) h = (h ^ static_cast<uint32_t>(v)) * 1099511628211ull;
    k += "|" + std::to_string(offs->size()) + "|" + std::to_string(h);
  }
  return k;
}

int run_plan(ocp_ctx c, const float* pts, int f3, const Plan& p,
             const std::string& key_in) {
  const std::string key = key_in + (p.permutation ? "|perm" : "");
  if (c->graph_mode) {
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
      hipGraph_t g = nullptr;
      OCP_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const bool ok = enqueue_pipeline(c, pts, f3 != 0, p, true);
      hipError_t e = hipGetLastError();
      hipError_t e2 = hipStreamEndCapture(c->stream, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(c, OCP_ERR_HIP, std::string("graph capture: ") +
                                        hipGetErrorString(e != hipSuccess ? e : e2));
      }
      if (!ok) {
        if (g) (void)hipGraphDestroy(g);
        return fail(c, OCP_ERR_INVALID, "octree stage: a stage array is missing");
      }
      hipGraphExec_t ex = nullptr;
      e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(c, OCP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      it = c->graphs.emplace(key, ex).first;
    }
    OCP_HIP(c, hipGraphLaunch(it->second, c->stream));
    c->ev_recorded = false;
  } else {
    c->order_err = hipSuccess;
    const bool ok = enqueue_pipeline(c, pts, f3 != 0, p, false);
    OCP_HIP(c, c->order_err);
    OCP_CHECK_LAUNCH(c);
    if (!ok) return fail(c, OCP_ERR_INVALID, "octree stage: a stage array is missing");
    c->ev_recorded = c->stage_timing;
  }
  c->last_plan = p;
  c->have_result = true;
  return OCP_OK;
}

// A lookback wait that gave up since the last check: the word is cleared
// (the next launch runs normally), the pending build result is dropped and
// the error is returned. `hc` holds the control block read after the
// stream was synchronised.
int report_timeout(ocp_ctx c, const uint32_t* hc) {
  if (hc[kCtrlTimeout] == 0) return OCP_OK;
  c->have_result = false;
  launch_fill_u32(c->ctrl + kCtrlTimeout, 0u, 1, c->stream);
  OCP_CHECK_LAUNCH(c);
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return fail(c, OCP_ERR_STATE,
              "a lookback wait gave up (wait limit reached); the result of the work "
              "enqueued since the last check is invalid");
}

// Synchronises the stream and reports a wait that gave up.
int sync_and_check(ocp_ctx c) {
  uint32_t hc[kCtrlWords];
  OCP_HIP(c, hipMemcpyAsync(hc, c->ctrl, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return report_timeout(c, hc);
}

// After a build: if the octree did not fit, grow the node buffer and redo
// stage 7 (its inputs are intact). Synchronises.
int settle(ocp_ctx c) {
  if (!c->have_result) return fail(c, OCP_ERR_STATE, "no build result (none has run, or it failed)");
  uint32_t hc[kCtrlWords];
  OCP_HIP(c, hipMemcpyAsync(hc, c->ctrl, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  if (int r = report_timeout(c, hc)) return r;
  if (hc[kCtrlOctOk] == 0) {
    const int64_t need = hc[kCtrlOctTotal];
    clear_graphs(c);
    if (int r = alloc_nodes(c, need + need / 4 + 1024)) return r;
    if (!enqueue_octree(c, c->last_plan, true)) {
      c->have_result = false;
      return fail(c, OCP_ERR_INVALID, "octree stage: a stage array is missing");
    }
    OCP_CHECK_LAUNCH(c);
    OCP_HIP(c, hipMemcpyAsync(hc, c->ctrl, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
    OCP_HIP(c, hipStreamSynchronize(c->stream));
    if (hc[kCtrlOctOk] == 0) return fail(c, OCP_ERR_CAPACITY, "octree node buffer");
  }
  return OCP_OK;
}

int check_ctx(ocp_ctx c) { return c ? OCP_OK : OCP_ERR_INVALID; }

void free_learn(LearnState& L) {
  dev_free(L.model); dev_free(L.meta); dev_free(L.kleaf); dev_free(L.fb); dev_free(L.loss);
  dev_free(L.ctab);
  dev_free(L.ring); dev_free(L.samp_a); dev_free(L.samp_b);
  LearnTables& t = L.tb;
  dev_free(t.rows); dev_free(t.hist); dev_free(t.bstart); dev_free(t.hist_a);
  dev_free(t.hist_b); dev_free(t.seg_start); dev_free(t.tile_start); dev_free(t.tiles_b);
  dev_free(t.big_start); dev_free(t.big_end); dev_free(t.big_tiles); dev_free(t.info);
  dev_free(t.big_hist); dev_free(t.cursor); dev_free(t.big_cursor);
  if (L.h_info) (void)hipHostFree(L.h_info);
  if (L.ev_info) (void)hipEventDestroy(L.ev_info);
  L = LearnState{};
}

}  // namespace

// ======================================================================
extern "C" {

int ocp_create(ocp_ctx* out, int device, void* stream) {
  if (!out) return OCP_ERR_INVALID;
  *out = nullptr;
  auto* c = new ocp_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return OCP_ERR_HIP;
  }
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return OCP_ERR_HIP;
    }
    c->own_stream = true;
  }
  if (hipMalloc(reinterpret_cast<void**>(&c->ctrl), sizeof(uint32_t) * kCtrlWords) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&c->s_hist), sizeof(uint32_t) * 1024) != hipSuccess ||
      hipMemsetAsync(c->ctrl, 0, sizeof(uint32_t) * kCtrlWords, c->stream) != hipSuccess) {
    ocp_destroy(c);
    return OCP_ERR_HIP;
  }
  std::string why;
  if (int r = scratch_init(device, c->stream, &why)) {
    ocp_destroy(c);
    return r;
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  for (auto& e : c->kev) (void)hipEventCreate(&e);
  *out = c;
  return OCP_OK;
}

int ocp_destroy(ocp_ctx c) {
  if (!c) return OCP_ERR_INVALID;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  clear_graphs(c);
  free_workspace(c->ws);
  dev_free(c->ctrl);
  dev_free(c->s_flags);
  dev_free(c->s_hist);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->kev)
    if (e) (void)hipEventDestroy(e);
  free_learn(c->learn);
  dev_free(c->cold_buf);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return OCP_OK;
}

const char* ocp_last_error(ocp_ctx c) { return c ? c->err.c_str() : "null context"; }

int ocp_synchronize(ocp_ctx c) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  return sync_and_check(c);
}

int ocp_set_wait_limit(ocp_ctx c, int64_t polls) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (polls > int64_t(0xfffffffe)) return fail(c, OCP_ERR_INVALID, "polls > 2^32-2");
  const uint32_t v = polls < 0 ? 0u : static_cast<uint32_t>(polls) + 1u;
  launch_fill_u32(c->ctrl + kCtrlWaitLimit, v, 1, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_malloc(ocp_ctx c, void** ptr, size_t bytes) {
  if (check_ctx(c) || !ptr) return OCP_ERR_INVALID;
  OCP_HIP(c, hipMalloc(ptr, bytes ? bytes : 1));
  return OCP_OK;
}
int ocp_device_memory(ocp_ctx c, int64_t* free_bytes, int64_t* total_bytes) {
  if (check_ctx(c) || !free_bytes || !total_bytes) return OCP_ERR_INVALID;
  OCP_HIP(c, hipSetDevice(c->device));
  size_t fr = 0, tot = 0;
  OCP_HIP(c, hipMemGetInfo(&fr, &tot));
  *free_bytes = static_cast<int64_t>(fr);
  *total_bytes = static_cast<int64_t>(tot);
  return OCP_OK;
}
int ocp_free(ocp_ctx c, void* ptr) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  if (ptr) OCP_HIP(c, hipFree(ptr));
  return OCP_OK;
}
int ocp_memcpy_h2d(ocp_ctx c, void* dst, const void* src, size_t bytes) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!bytes) return OCP_OK;
  OCP_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}
int ocp_memcpy_d2h(ocp_ctx c, void* dst, const void* src, size_t bytes) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!bytes) return OCP_OK;
  OCP_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}
int ocp_copy(ocp_ctx c, void* dst, const void* src, size_t bytes) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!bytes) return OCP_OK;
  if (!dst || !src || (bytes & 15u) || (reinterpret_cast<uintptr_t>(dst) & 15u) ||
      (reinterpret_cast<uintptr_t>(src) & 15u))
    return fail(c, OCP_ERR_INVALID, "ocp_copy needs 16-byte aligned pointers and size");
  launch_copy(dst, src, static_cast<int64_t>(bytes), c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_memset(ocp_ctx c, void* dst, int value, size_t bytes) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!bytes) return OCP_OK;
  OCP_HIP(c, hipMemsetAsync(dst, value, bytes, c->stream));
  return OCP_OK;
}

// ---------------------------------------------------------------- stages
int ocp_k_InitRandomVec4(ocp_ctx c, float* data, int32_t n, float mn, float rg, int32_t seed) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 0) return OCP_OK;
  if (!data) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (reinterpret_cast<uintptr_t>(data) & 15u)
    return fail(c, OCP_ERR_INVALID, "float4 output must be 16-byte aligned");
  launch_init_random(data, n, mn, rg, seed, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

static int morton_common(ocp_ctx c, const float* data, uint32_t* keys, size_t n,
                         float mn, float rg, bool f3) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n == 0) return OCP_OK;
  if (!data || !keys) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (static_cast<int64_t>(n) > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  if (!f3 && (reinterpret_cast<uintptr_t>(data) & 15u))
    return fail(c, OCP_ERR_INVALID, "float4 input must be 16-byte aligned");
  const SegView sv = single_view(static_cast<int64_t>(n), mn, rg);
  launch_morton(data, f3, keys, sv, nullptr,
                static_cast<int32_t>(cdiv(static_cast<int64_t>(n), morton_tile(c->tune))),
                nullptr, c->tune, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_ComputeMortonCode(ocp_ctx c, const float* data, uint32_t* keys, size_t n,
                            float mn, float rg) {
  return morton_common(c, data, keys, n, mn, rg, false);
}
int ocp_k_ComputeMortonCode3(ocp_ctx c, const float* data, uint32_t* keys, size_t n,
                             float mn, float rg) {
  return morton_common(c, data, keys, n, mn, rg, true);
}

int ocp_k_GlobalHistogram(ocp_ctx c, const uint32_t* sort, uint32_t* hist, uint32_t size) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!hist || (size && !sort)) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (size > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "size > 2^30-1");
  launch_fill_u32(hist, 0u, 1024, c->stream);
  const SegView sv = single_view(size, 0.0f, 1.0f);
  launch_histogram(sort, sv, nullptr, static_cast<int32_t>(cdiv(size, morton_tile(c->tune))),
                   hist, c->tune, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_DigitBinning(ocp_ctx c, const uint32_t* hist, const uint32_t* sort, uint32_t* alt,
                       uint32_t size, uint32_t radix_shift) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (size == 0) return OCP_OK;
  if (!hist || !sort || !alt) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (radix_shift > 24 || (radix_shift & 7u)) return fail(c, OCP_ERR_INVALID, "radix_shift must be 0, 8, 16 or 24");
  if (size > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "size > 2^30-1");
  const int32_t tiles = static_cast<int32_t>(cdiv(size, sort_tile(c->tune)));
  if (int r = ensure_scratch(c, int64_t(tiles) * 256)) return r;
  launch_begin(c->ctrl, nullptr, 0, c->stream);
  launch_onesweep_pass(sort, alt, single_view(size, 0.0f, 1.0f), nullptr, tiles, hist,
                       c->s_flags, c->ctrl, static_cast<int>(radix_shift >> 3),
                       /*stable=*/true, c->tune, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_SortKeys(ocp_ctx c, uint32_t* keys, uint32_t* alt, uint32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 1) return OCP_OK;
  if (!keys || !alt) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  const int32_t tiles = static_cast<int32_t>(cdiv(n, sort_tile(c->tune)));
  if (int r = ensure_scratch(c, int64_t(tiles) * 256)) return r;
  const SegView sv = single_view(n, 0.0f, 1.0f);
  launch_begin(c->ctrl, c->s_hist, 1024, c->stream);
  launch_histogram(keys, sv, nullptr, static_cast<int32_t>(cdiv(n, morton_tile(c->tune))),
                   c->s_hist, c->tune, c->stream);
  uint32_t* src = keys;
  uint32_t* dst = alt;
  for (int pass = 0; pass < 4; ++pass) {
    launch_onesweep_pass(src, dst, sv, nullptr, tiles, c->s_hist, c->s_flags, c->ctrl, pass,
                         pass != 0, c->tune, c->stream);
    std::swap(src, dst);
  }
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_SortKeys_static(ocp_ctx c, uint32_t* keys, uint32_t* alt, uint32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 1) return OCP_OK;
  if (!keys || !alt) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n > kStaticSortMaxKeys)
    return fail(c, OCP_ERR_CAPACITY, "ocp_k_SortKeys_static sorts at most 2^24 keys");
  const int32_t tiles = static_cast<int32_t>(cdiv(n, sort_tile(c->tune)));
  if (int64_t(tiles) * 256 > kStaticFlagWords)
    return fail(c, OCP_ERR_CAPACITY, "ocp_k_SortKeys_static: too many sort tiles");
  uint8_t* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!base) return fail(c, OCP_ERR_STATE, "static sort scratch not set up for this device");
  uint32_t* dev_epoch = reinterpret_cast<uint32_t*>(base);
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + kStaticHeaderBytes);
  uint64_t* flags = reinterpret_cast<uint64_t*>(base + kStaticHeaderBytes + kStaticHistBytes);
  const SegView sv = single_view(n, 0.0f, 1.0f);
  launch_begin_shared(c->ctrl, dev_epoch, hist, 1024, c->stream);
  launch_histogram(keys, sv, nullptr, static_cast<int32_t>(cdiv(n, morton_tile(c->tune))), hist,
                   c->tune, c->stream);
  uint32_t* src = keys;
  uint32_t* dst = alt;
  for (int pass = 0; pass < 4; ++pass) {
    launch_onesweep_pass(src, dst, sv, nullptr, tiles, hist, flags, c->ctrl, pass, pass != 0,
                         c->tune, c->stream);
    std::swap(src, dst);
  }
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_SortPairs(ocp_ctx c, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                    uint32_t* vals_alt, uint32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 1) return OCP_OK;
  if (!keys || !vals || !keys_alt || !vals_alt) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  const int32_t tiles = static_cast<int32_t>(cdiv(n, sort_tile(c->tune)));
  if (int r = ensure_scratch(c, int64_t(tiles) * 256)) return r;
  const SegView sv = single_view(n, 0.0f, 1.0f);
  launch_begin(c->ctrl, c->s_hist, 1024, c->stream);
  launch_histogram(keys, sv, nullptr, static_cast<int32_t>(cdiv(n, morton_tile(c->tune))),
                   c->s_hist, c->tune, c->stream);
  uint32_t* src = keys;
  uint32_t* dst = keys_alt;
  uint32_t* vs = vals;
  uint32_t* vd = vals_alt;
  for (int pass = 0; pass < 4; ++pass) {
    launch_onesweep_pass_kv(src, dst, vs, vd, sv, nullptr, tiles, c->s_hist, c->s_flags, c->ctrl,
                            pass, c->tune, c->stream);
    std::swap(src, dst);
    std::swap(vs, vd);
  }
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_CountUniquePairs(ocp_ctx c, uint32_t* keys, uint32_t* vals, int32_t* num_unique_out,
                           int32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!num_unique_out || (n > 0 && (!keys || !vals))) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n < 0) return fail(c, OCP_ERR_INVALID, "n < 0");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  launch_fill_u32(reinterpret_cast<uint32_t*>(num_unique_out), 0u, 1, c->stream);
  if (n == 0) {
    OCP_CHECK_LAUNCH(c);
    return OCP_OK;
  }
  const int32_t tiles = static_cast<int32_t>(cdiv(n, scan_tile(c->tune)));
  if (int r = ensure_scratch(c, 2 * int64_t(tiles))) return r;
  launch_begin(c->ctrl, nullptr, 0, c->stream);
  launch_unique(keys, keys, single_view(n, 0.0f, 1.0f), nullptr, tiles, c->s_flags, c->ctrl,
                num_unique_out, nullptr, c->tune, c->stream, vals, vals);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_CountUnique(ocp_ctx c, uint32_t* keys, int32_t* num_unique_out, int32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (!num_unique_out || (n > 0 && !keys)) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n < 0) return fail(c, OCP_ERR_INVALID, "n < 0");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  launch_fill_u32(reinterpret_cast<uint32_t*>(num_unique_out), 0u, 1, c->stream);
  if (n == 0) {
    OCP_CHECK_LAUNCH(c);
    return OCP_OK;
  }
  const int32_t tiles = static_cast<int32_t>(cdiv(n, scan_tile(c->tune)));
  if (int r = ensure_scratch(c, 2 * int64_t(tiles))) return r;  // tile flags + group sums
  launch_begin(c->ctrl, nullptr, 0, c->stream);
  launch_unique(keys, keys, single_view(n, 0.0f, 1.0f), nullptr, tiles, c->s_flags, c->ctrl,
                num_unique_out, nullptr, c->tune, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_BuildRadixTree(ocp_ctx c, int32_t n, const uint32_t* codes, uint8_t* prefix_n,
                         uint8_t* hl, uint8_t* hr, int32_t* lc, int32_t* parent) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 0) return OCP_OK;
  if (!codes || !prefix_n || !hl || !hr || !lc || !parent)
    return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  launch_radix_tree(codes, prefix_n, hl, hr, lc, parent, nullptr, single_view(n + 1, 0.0f, 1.0f),
                    nullptr, static_cast<int32_t>(cdiv(n, kNodeTile)), nullptr, n, c->ctrl,
                    c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_EdgeCount(ocp_ctx c, const uint8_t* prefix_n, const int32_t* parents,
                    int32_t* edge_count, int32_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 0) return OCP_OK;
  if (!prefix_n || !parents || !edge_count) return fail(c, OCP_ERR_INVALID, "null pointer");
  launch_edge_scan(prefix_n, parents, nullptr, edge_count, nullptr, single_view(n, 0.0f, 1.0f),
                   nullptr, static_cast<int32_t>(cdiv(n, scan_tile(c->tune))), nullptr, n,
                   nullptr, c->ctrl, nullptr, 1, c->tune, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_PartialSum(ocp_ctx c, const int32_t* data, int64_t first, int64_t last,
                     int32_t* d_first) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  const int64_t n = last - first;
  if (n <= 0) return OCP_OK;
  if (!data || !d_first || first < 0) return fail(c, OCP_ERR_INVALID, "bad arguments");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  const int32_t tiles = static_cast<int32_t>(cdiv(n, scan_tile(c->tune)));
  if (int r = ensure_scratch(c, 2 * int64_t(tiles))) return r;  // tile flags + group sums
  launch_begin(c->ctrl, nullptr, 0, c->stream);
  launch_edge_scan(nullptr, nullptr, nullptr, const_cast<int32_t*>(data + first), d_first,
                   single_view(n, 0.0f, 1.0f), nullptr, tiles, nullptr,
                   static_cast<int32_t>(n), c->s_flags, c->ctrl, nullptr, 2, c->tune,
                   c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_InitOctNodes(ocp_ctx c, ocp_oct_node* nodes, int32_t n_alloc, const uint32_t* codes,
                       const uint8_t* prefix_n, int32_t n_brt, float mn, float rg) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n_alloc <= 0) return OCP_OK;
  if (!nodes || !codes || (n_brt > 0 && !prefix_n)) return fail(c, OCP_ERR_INVALID, "null pointer");
  launch_fill_u32(reinterpret_cast<uint32_t*>(nodes), 0u,
                  int64_t(n_alloc) * (sizeof(OctNode) / 4), c->stream);
  launch_oct_root(reinterpret_cast<OctNode*>(nodes), codes, prefix_n, n_brt, mn, rg, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

static OctArgs stage_oct_args(ocp_oct_node* nodes, const int32_t* offs, const int32_t* counts,
                              const uint32_t* codes, const uint8_t* pn, const uint8_t* hl,
                              const uint8_t* hr, const int32_t* par, const int32_t* lc,
                              int32_t n, int32_t n_alloc) {
  OctArgs a{};
  a.nodes = reinterpret_cast<OctNode*>(nodes);
  a.k_host = n_alloc;
  a.codes = codes;
  a.prefix_n = pn;
  a.has_leaf_left = hl;
  a.has_leaf_right = hr;
  a.left_child = lc;
  a.parent = par;
  a.edge = counts;
  a.offsets = offs;
  a.n_brt_host = n;
  return a;
}

int ocp_k_MakeOctNodes(ocp_ctx c, ocp_oct_node* nodes, const int32_t* offs,
                       const int32_t* counts, const uint32_t* codes, const uint8_t* pn,
                       const int32_t* par, float mn, float rg, int32_t n, int32_t n_alloc) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 0 || n_alloc <= 0) return OCP_OK;
  if (!nodes || !offs || !counts || !codes || !pn || !par)
    return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n >= kMaxSegPoints) return fail(c, OCP_ERR_CAPACITY, "N >= 2^27");
  const OctArgs a = stage_oct_args(nodes, offs, counts, codes, pn, nullptr, nullptr, par,
                                   nullptr, n, n_alloc);
  const SegView sv = single_view(n + 1, mn, rg);
  const int32_t tiles = static_cast<int32_t>(cdiv(n, kNodeTile));
  // No children arrays in this call: the push-style kernels.
  launch_oct_scatter(a, sv, nullptr, tiles, 1, c->stream);
  launch_oct_finalize(a, sv, nullptr, tiles, 1 | 4, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_k_LinkLeafNodes(ocp_ctx c, ocp_oct_node* nodes, const int32_t* offs,
                        const int32_t* counts, const uint32_t* codes, const uint8_t* hl,
                        const uint8_t* hr, const uint8_t* pn, const int32_t* par,
                        const int32_t* lc, int32_t n, int32_t n_alloc) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n <= 0 || n_alloc <= 0) return OCP_OK;
  if (!nodes || !offs || !counts || !codes || !hl || !hr || !pn || !par || !lc)
    return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n >= kMaxSegPoints) return fail(c, OCP_ERR_CAPACITY, "N >= 2^27");
  const OctArgs a = stage_oct_args(nodes, offs, counts, codes, pn, hl, hr, par, lc, n, n_alloc);
  const SegView sv = single_view(n + 1, 0.0f, 1.0f);
  const int32_t tiles = static_cast<int32_t>(cdiv(n, kNodeTile));
  launch_oct_scatter(a, sv, nullptr, tiles, 2 | 4, c->stream);
  launch_oct_finalize(a, sv, nullptr, tiles, 2, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

// ---------------------------------------------------------------- pipeline
int ocp_build(ocp_ctx c, const float* points, int float3, int64_t n, float mn, float rg) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n < 0) return fail(c, OCP_ERR_INVALID, "n < 0");
  if (n > 0 && !points) return fail(c, OCP_ERR_INVALID, "null points");
  if (n > kMaxSegPoints) return fail(c, OCP_ERR_CAPACITY, "a single cloud holds at most 2^27 points");
  if (!(rg > 0.0f)) return fail(c, OCP_ERR_INVALID, "range must be > 0");
  if (!float3 && (reinterpret_cast<uintptr_t>(points) & 15u))
    return fail(c, OCP_ERR_INVALID, "float4 input must be 16-byte aligned");
  const TuneTable& tt = c->tune;
  if (int r = ensure_workspace(c, n, 1, flag_words_for(tt, n, 1))) return r;
  Plan p = plain_plan(tt, n, mn, rg, c->permutation);
  if (int r = ensure_perm(c)) return r;
  c->last.valid = true;
  c->last.batched = false;
  c->last.pts = points;
  c->last.float3 = float3;
  c->last.n = n;
  c->last.mn = mn;
  c->last.rg = rg;
  return run_plan(c, points, float3, p, graph_key("plain", points, float3, n, mn, rg, nullptr, tt));
}

int ocp_build_autobox(ocp_ctx c, const float* points, int float3, int64_t n) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n < 0) return fail(c, OCP_ERR_INVALID, "n < 0");
  if (n > 0 && !points) return fail(c, OCP_ERR_INVALID, "null points");
  if (n > kMaxSegPoints) return fail(c, OCP_ERR_CAPACITY, "a single cloud holds at most 2^27 points");
  if (!float3 && (reinterpret_cast<uintptr_t>(points) & 15u))
    return fail(c, OCP_ERR_INVALID, "float4 input must be 16-byte aligned");
  const TuneTable& tt = c->tune;
  if (int r = ensure_workspace(c, n, 1, flag_words_for(tt, n, 1))) return r;
  Plan p = plain_plan(tt, n, 0.0f, 1.0f, c->permutation);
  p.device_boxes = true;
  p.permutation = c->permutation;
  if (int r = ensure_perm(c)) return r;
  p.sv.mn = c->ws.boxes;
  p.sv.rg = c->ws.boxes + c->ws.seg_cap;
  c->last.valid = true;
  c->last.batched = false;
  c->last.pts = points;
  c->last.float3 = float3;
  c->last.n = n;
  return run_plan(c, points, float3, p,
                  graph_key("plain-autobox", points, float3, n, 0, 0, nullptr, tt));
}

int ocp_transform_points(ocp_ctx c, const float* src, float* dst, int float3, int64_t n,
                         const float* affine, float lo, float hi) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n < 0) return fail(c, OCP_ERR_INVALID, "n < 0");
  if (!affine) return fail(c, OCP_ERR_INVALID, "null affine");
  if (n > 0 && (!src || !dst)) return fail(c, OCP_ERR_INVALID, "null points");
  if (!(lo <= hi)) return fail(c, OCP_ERR_INVALID, "lo > hi");
  Affine m;
  std::memcpy(m.m, affine, sizeof(m.m));
  launch_transform(src, dst, n, float3 != 0, m, lo, hi, c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_build_batched(ocp_ctx c, const float* points, int float3, const int32_t* offsets,
                      const float* smin, const float* srg, int32_t n_seg) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  Plan p;
  if (int r = batched_plan(c, offsets, smin, srg, n_seg, c->permutation, &p)) return r;
  if (p.n_total > 0 && !points) return fail(c, OCP_ERR_INVALID, "null points");
  p.permutation = c->permutation;
  if (int r = ensure_perm(c)) return r;
  if (!float3 && (reinterpret_cast<uintptr_t>(points) & 15u))
    return fail(c, OCP_ERR_INVALID, "float4 input must be 16-byte aligned");
  c->last.valid = true;
  c->last.batched = true;
  c->last.pts = points;
  c->last.float3 = float3;
  c->last.offsets.assign(offsets, offsets + n_seg + 1);
  if (smin) {
    c->last.seg_min.assign(smin, smin + n_seg);
    c->last.seg_range.assign(srg, srg + n_seg);
  } else {
    c->last.seg_min.clear();
    c->last.seg_range.clear();
  }
  return run_plan(c, points, float3, p,
                  graph_key(smin ? "batch" : "batch-autobox", points, float3, p.n_total, 0, 0,
                            &c->last.offsets, c->tune));
}

int ocp_workspace_size(const ocp_tune* tune, int64_t n, int32_t n_seg, int permutation,
                       ocp_workspace_sizes* out) {
  if (!out || n < 0 || n > kMaxPoints || n_seg < 0) return OCP_ERR_INVALID;
  TuneTable tt;
  if (tune) {
    tt.sort_waves = tune->sort_waves;
    tt.sort_kpt = tune->sort_kpt;
    tt.morton_ppt = tune->morton_ppt;
    tt.morton_grid = tune->morton_grid;
    tt.scan_ipt = tune->scan_ipt;
    tt.sort_bits = tune->sort_bits;
    tt.sort10_kpt = tune->sort10_kpt;
    tt.oct_order = tune->oct_order;
    tt.sort_msd = tune->sort_msd;
    if (!sort_config_supported(tt.sort_waves, tt.sort_kpt) && !(tt.sort_waves == 8 && tt.sort_kpt == 15))
      return OCP_ERR_INVALID;
    if (!scan_config_supported(tt.scan_ipt) || !sort10_config_supported(tt.sort10_kpt) ||
        (tt.oct_order != 0 && tt.oct_order != 1) || (tt.sort_msd != 0 && tt.sort_msd != 1))
      return OCP_ERR_INVALID;
  }
  // The allocations of ensure_workspace, alloc_nodes and ensure_perm for a
  // first build of this shape, plus the context's control words and digit
  // counts (ocp_create).
  const int64_t cap = std::max<int64_t>(n, 4096);
  const int64_t seg = std::max<int32_t>(n_seg, 1);
  int64_t b = 2 * 4 * cap          // keys_a, keys_b
              + 4 * 1 * cap        // prefix_n, hl, hr, parent prefix lengths
              + 4 * 4 * cap        // left child, parent, edge counts, offsets
              + cap * int64_t(sizeof(OctNode)) + 16;  // nodes (K_cap starts at N)
  b += 4 * (seg * (hist_stride(10) + kCursorWords) + 2 * seg + (seg + 1) + 6 * (seg + 1) +
            4 * seg);
  b += 8 * std::max<int64_t>(flag_words_for(tt, n, static_cast<int32_t>(seg)), 1 << 16);
  if (permutation) b += 2 * 4 * cap;
  b += 4 * (kCtrlWords + 1024);
  out->build_bytes = b;
  const int64_t sort_tiles = cdiv(n, sort_tile(tt));
  const int64_t scan_tiles = cdiv(n, scan_tile(tt));
  out->sort_bytes = 8 * std::max<int64_t>(sort_tiles * 256, 1 << 16);
  out->unique_bytes = 8 * std::max<int64_t>(2 * scan_tiles, 1 << 16);
  out->scan_bytes = out->unique_bytes;
  return OCP_OK;
}

int ocp_check_batch(const int32_t* offsets, const float* smin, const float* srg,
                    int32_t n_seg) {
  std::string why;
  return validate_batch(offsets, smin, srg, n_seg, &why);
}

int ocp_set_permutation(ocp_ctx c, int enable) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  c->permutation = enable != 0;
  return OCP_OK;
}

int ocp_set_graph_mode(ocp_ctx c, int enable) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  c->graph_mode = enable != 0;
  return OCP_OK;
}

int ocp_get_results(ocp_ctx c, ocp_results* out) {
  if (check_ctx(c) || !out) return OCP_ERR_INVALID;
  if (int r = settle(c)) return r;
  const Workspace& w = c->ws;
  out->sorted_keys = w.keys_a;
  out->unique_keys = w.keys_b;
  out->prefix_n = w.prefix_n;
  out->has_leaf_left = w.hl;
  out->has_leaf_right = w.hr;
  out->left_child = w.lc;
  out->parent = w.par;
  out->edge_count = w.edge;
  out->offsets = w.offs;
  out->nodes = reinterpret_cast<const ocp_oct_node*>(w.nodes);
  out->n_unique = w.counts;
  out->n_oct = w.counts + w.seg_cap;
  out->oct_start = w.oct_start;
  out->n_seg = c->last_plan.n_seg;
  out->n_points = c->last_plan.n_total;
  const bool boxes = c->last_plan.device_boxes || c->last_plan.sv.start;
  out->seg_min = boxes ? w.boxes : nullptr;
  out->seg_range = boxes ? w.boxes + w.seg_cap : nullptr;
  out->sorted_index = c->last_plan.permutation ? w.vals_a : nullptr;
  out->unique_index = c->last_plan.permutation ? w.vals_b : nullptr;
  return OCP_OK;
}

int ocp_get_counts(ocp_ctx c, int32_t* n_unique, int32_t* n_oct) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (int r = settle(c)) return r;
  const Workspace& w = c->ws;
  const int32_t ns = c->last_plan.n_seg;
  if (n_unique)
    OCP_HIP(c, hipMemcpy(n_unique, w.counts, sizeof(int32_t) * ns, hipMemcpyDeviceToHost));
  if (n_oct)
    OCP_HIP(c, hipMemcpy(n_oct, w.counts + w.seg_cap, sizeof(int32_t) * ns, hipMemcpyDeviceToHost));
  return OCP_OK;
}

extern "C++" {
namespace {

// Device memory of one format call, freed when the call returns (after the
// stream is synchronised).
struct TempBuffers {
  std::vector<void*> ptrs;
  ~TempBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(ocp_ctx c, T** p, int64_t n) {
    if (int r = dev_alloc(c, p, n)) return r;
    ptrs.push_back(*p);
    return OCP_OK;
  }
};

// Total slots of the last build and each segment's root level (validator
// and compact form). Settles the build first.
int format_setup(ocp_ctx c, TempBuffers& tb, int64_t* k_total, int32_t** root_level,
                 uint32_t** root_code) {
  if (int r = settle(c)) return r;
  const Workspace& w = c->ws;
  const Plan& p = c->last_plan;
  int32_t kt = 0;
  OCP_HIP(c, hipMemcpyAsync(&kt, w.oct_start + p.n_seg, sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  *k_total = kt;
  if (int r = tb.alloc(c, root_level, p.n_seg)) return r;
  if (int r = tb.alloc(c, root_code, p.n_seg)) return r;
  launch_seg_roots(w.keys_b, w.prefix_n, w.counts, p.sv.start, p.n_seg, *root_level, *root_code,
                   c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

}  // namespace
}  // extern "C++"

int ocp_validate_octree(ocp_ctx c, ocp_tree_report* out) {
  if (check_ctx(c) || !out) return OCP_ERR_INVALID;
  TempBuffers tb;
  int64_t kt = 0;
  int32_t* root_level = nullptr;
  uint32_t* root_code = nullptr;
  if (int r = format_setup(c, tb, &kt, &root_level, &root_code)) return r;
  const Workspace& w = c->ws;
  const Plan& p = c->last_plan;
  uint8_t* cov = nullptr;
  unsigned long long* rep = nullptr;
  if (int r = tb.alloc(c, &cov, kt)) return r;
  if (int r = tb.alloc(c, &rep, kTreeReportFields)) return r;
  OCP_HIP(c, hipMemsetAsync(cov, 0, static_cast<size_t>(std::max<int64_t>(kt, 1)), c->stream));
  OCP_HIP(c, hipMemsetAsync(rep, 0, sizeof(unsigned long long) * kTreeReportFields, c->stream));
  launch_validate(w.nodes, w.edge, w.offs, w.counts, w.oct_start, p.sv.start, p.n_seg, p.sv.mn,
                  p.sv.rg, p.sv.min1, p.sv.range1, root_level, root_code, cov, rep, c->stream);
  OCP_CHECK_LAUNCH(c);
  unsigned long long h[kTreeReportFields];
  OCP_HIP(c, hipMemcpyAsync(h, rep, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  static_assert(sizeof(ocp_tree_report) == kTreeReportFields * sizeof(int64_t),
                "report fields");
  int64_t* o = reinterpret_cast<int64_t*>(out);
  for (int f = 0; f < kTreeReportFields; ++f) o[f] = static_cast<int64_t>(h[f]);
  return OCP_OK;
}

int ocp_compact_octree(ocp_ctx c, ocp_compact_node* nodes, int64_t node_cap, int32_t* children,
                       int64_t child_cap, int64_t* n_nodes, int64_t* n_children) {
  if (check_ctx(c) || !n_nodes || !n_children || node_cap < 0 || child_cap < 0)
    return OCP_ERR_INVALID;
  TempBuffers tb;
  int64_t kt = 0;
  int32_t* root_level = nullptr;
  uint32_t* root_code = nullptr;
  if (int r = format_setup(c, tb, &kt, &root_level, &root_code)) return r;
  const Workspace& w = c->ws;
  const Plan& p = c->last_plan;
  int32_t* cnt = nullptr;
  uint32_t* bsum = nullptr;
  unsigned long long* bad = nullptr;
  const int64_t nb = compact_scan_blocks(kt);
  if (int r = tb.alloc(c, &cnt, kt)) return r;
  if (int r = tb.alloc(c, &bsum, nb + 1)) return r;
  if (int r = tb.alloc(c, &bad, 1)) return r;
  OCP_HIP(c, hipMemsetAsync(bad, 0, sizeof(unsigned long long), c->stream));
  launch_compact_count(w.nodes, w.oct_start, p.n_seg, p.sv.mn, p.sv.rg, p.sv.min1, p.sv.range1,
                       root_level, kt, cnt, bsum, bad, c->stream);
  OCP_CHECK_LAUNCH(c);
  uint32_t total = 0;
  unsigned long long hbad = 0;
  OCP_HIP(c, hipMemcpyAsync(&total, bsum + nb, sizeof(total), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  *n_nodes = kt;
  *n_children = total;
  if (hbad)
    return fail(c, OCP_ERR_STATE, std::to_string(hbad) +
                                      " octree slots have no exact compact form (corner off "
                                      "its level's grid or cell size not range / 2^k)");
  if (!nodes || !children) return OCP_OK;  // size query
  if (node_cap < kt || child_cap < static_cast<int64_t>(total))
    return fail(c, OCP_ERR_CAPACITY, "compact octree: output arrays too small (need " +
                                         std::to_string(kt) + " nodes, " +
                                         std::to_string(total) + " children)");
  launch_compact_write(w.nodes, w.oct_start, p.n_seg, p.sv.mn, p.sv.rg, p.sv.min1, p.sv.range1,
                       root_level, cnt, nodes, children, c->stream);
  OCP_CHECK_LAUNCH(c);
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}

int ocp_expand_octree(ocp_ctx c, const ocp_compact_node* nodes, const int32_t* children,
                      int64_t n_children, const int32_t* oct_start, int32_t n_seg,
                      const float* seg_min, const float* seg_range, float min_coord, float range,
                      ocp_oct_node* out) {
  if (check_ctx(c) || !nodes || !oct_start || !out || n_seg < 1 || n_children < 0 ||
      (n_children > 0 && !children) || (!seg_min) != (!seg_range))
    return fail(c, OCP_ERR_INVALID, "expand octree: bad arguments");
  launch_expand(nodes, children, n_children, oct_start, n_seg, seg_min, seg_range, min_coord,
                range, reinterpret_cast<OctNode*>(out), c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_set_cold_timing(ocp_ctx c, int64_t flush_bytes) {
  if (check_ctx(c) || flush_bytes < 0) return OCP_ERR_INVALID;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(c->cold_buf);
  c->cold_words = 0;
  if (flush_bytes > 0) {
    const int64_t words = (flush_bytes + 3) / 4;
    if (int r = dev_alloc(c, &c->cold_buf, words)) return r;
    c->cold_words = words;
  }
  return OCP_OK;
}

int ocp_set_stage_timing(ocp_ctx c, int enable) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  c->stage_timing = enable != 0;
  return OCP_OK;
}

int ocp_get_stage_ms(ocp_ctx c, float* ms7) {
  if (check_ctx(c) || !ms7) return OCP_ERR_INVALID;
  if (!c->ev_recorded) return fail(c, OCP_ERR_STATE, "no eager build with stage timing on");
  OCP_HIP(c, hipEventSynchronize(c->ev[6]));
  for (int k = 0; k < 6; ++k) OCP_HIP(c, hipEventElapsedTime(&ms7[k], c->ev[k], c->ev[k + 1]));
  OCP_HIP(c, hipEventElapsedTime(&ms7[6], c->ev[0], c->ev[6]));
  return OCP_OK;
}

int ocp_get_kernel_ms(ocp_ctx c, float* ms, const char** names, int32_t max_kernels,
                      int32_t* n_kernels) {
  if (check_ctx(c) || !n_kernels) return OCP_ERR_INVALID;
  *n_kernels = kKernelCount;
  if (!c->ev_recorded) return fail(c, OCP_ERR_STATE, "no eager build with stage timing on");
  OCP_HIP(c, hipEventSynchronize(c->ev[6]));
  // Whether the MSD sort's LSD passes had work (read after the build ended).
  uint32_t msd_run = 0;
  if (c->kmsd && names)
    OCP_HIP(c, hipMemcpy(&msd_run, c->ctrl + kCtrlMsdRun, sizeof(msd_run), hipMemcpyDeviceToHost));
  for (int k = 0; k < kKernelCount && k < max_kernels; ++k) {
    if (names)
      names[k] = !(c->kmsd && k >= kKPass0 && k <= kKPass3) ? kKernelNames[k]
                 : (k == kKPass2 && msd_run == 0)           ? "msd_lsd_skipped"
                                                            : kMsdKernelNames[k - kKPass0];
    if (!ms) continue;
    ms[k] = 0.0f;  // a kernel the build did not launch (a fourth sort pass)
    if (c->kran[k]) OCP_HIP(c, hipEventElapsedTime(&ms[k], c->kev[2 * k], c->kev[2 * k + 1]));
  }
  return OCP_OK;
}

int ocp_time_builds(ocp_ctx c, const float* points, int float3, const int32_t* offsets,
                    const float* smin, const float* srg, int32_t n_seg, int warmup, int steps,
                    float* ms_per_build) {
  if (check_ctx(c) || !ms_per_build || steps < 1) return OCP_ERR_INVALID;
  // n_seg == 0: plain build of offsets[1] points with box (smin[0], srg[0]).
  auto once = [&]() -> int {
    if (n_seg == 0) return ocp_build(c, points, float3, offsets[1], smin[0], srg[0]);
    return ocp_build_batched(c, points, float3, offsets, smin, srg, n_seg);
  };
  if (int r = once()) return r;
  if (int r = settle(c)) return r;  // node capacity fixed before timing
  for (int k = 0; k < warmup; ++k)
    if (int r = once()) return r;
  hipEvent_t a, b;
  OCP_HIP(c, hipEventCreate(&a));
  OCP_HIP(c, hipEventCreate(&b));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  OCP_HIP(c, hipEventRecord(a, c->stream));
  for (int k = 0; k < steps; ++k)
    if (int r = once()) return r;
  OCP_HIP(c, hipEventRecord(b, c->stream));
  OCP_HIP(c, hipEventSynchronize(b));
  float ms = 0;
  OCP_HIP(c, hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms_per_build = ms / steps;
  return settle(c);
}

// ---------------------------------------------------------------- tuning
int ocp_get_tune(ocp_ctx c, ocp_tune* out) {
  if (check_ctx(c) || !out) return OCP_ERR_INVALID;
  out->sort_waves = c->tune.sort_waves;
  out->sort_kpt = c->tune.sort_kpt;
  out->morton_ppt = c->tune.morton_ppt;
  out->morton_grid = c->tune.morton_grid;
  out->scan_ipt = c->tune.scan_ipt;
  out->sort_bits = c->tune.sort_bits;
  out->sort10_kpt = c->tune.sort10_kpt;
  out->oct_order = c->tune.oct_order;
  out->sort_msd = c->tune.sort_msd;
  return OCP_OK;
}

int ocp_set_tune(ocp_ctx c, const ocp_tune* in) {
  if (check_ctx(c) || !in) return OCP_ERR_INVALID;
  const bool sort_ok = sort_config_supported(in->sort_waves, in->sort_kpt) ||
                       (in->sort_waves == 8 && in->sort_kpt == 15);
  if (!sort_ok || !morton_config_supported(in->morton_ppt) ||
      !scan_config_supported(in->scan_ipt) || in->morton_grid < 1 ||
      (in->sort_bits != 8 && in->sort_bits != 10) || !sort10_config_supported(in->sort10_kpt) ||
      (in->oct_order != 0 && in->oct_order != 1) || (in->sort_msd != 0 && in->sort_msd != 1))
    return fail(c, OCP_ERR_INVALID, "unsupported tile configuration");
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  c->tune.sort_waves = in->sort_waves;
  c->tune.sort_kpt = in->sort_kpt;
  c->tune.morton_ppt = in->morton_ppt;
  c->tune.morton_grid = in->morton_grid;
  c->tune.scan_ipt = in->scan_ipt;
  c->tune.sort_bits = in->sort_bits;
  c->tune.sort10_kpt = in->sort10_kpt;
  c->tune.oct_order = in->oct_order;
  c->tune.sort_msd = in->sort_msd;
  clear_graphs(c);
  c->tab_offsets.clear();
  return OCP_OK;
}

int ocp_autotune(ocp_ctx c, const float* points, int float3, int64_t n,
                 const int32_t* offsets, const float* smin, const float* srg, int32_t n_seg,
                 int reps, ocp_tune* chosen) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (reps < 1) reps = 3;
  std::vector<int32_t> plain_off = {0, static_cast<int32_t>(n)};
  const int32_t* offs = n_seg == 0 ? plain_off.data() : offsets;
  // A configuration's time: the best of three runs of `reps` builds (the
  // choices differ by fractions of a percent, about the spread of one run).
  auto time_cfg = [&](const ocp_tune& t, float* ms) -> int {
    if (int r = ocp_set_tune(c, &t)) return r;
    float best_run = 0;
    for (int run = 0; run < 3; ++run) {
      float m = 0;
      if (int r = ocp_time_builds(c, points, float3, offs, smin, srg, n_seg, run == 0 ? 1 : 0,
                                  reps, &m))
        return r;
      best_run = run == 0 ? m : std::min(best_run, m);
    }
    *ms = best_run;
    return OCP_OK;
  };
  ocp_tune best;
  ocp_get_tune(c, &best);
  float best_ms = 0;
  if (int r = time_cfg(best, &best_ms)) return r;
  // Coordinate descent, two rounds: sort tile of 8-bit digits, then the
  // pass count (four 8-bit passes, three 10-bit ones, or the MSD form: one
  // 10-bit scatter and LDS bucket sorts; two tile sizes each),
  // then scan tile, then Morton tile/grid. The second round revisits the
  // pass count under the chosen Morton shape: the Morton kernel's histogram
  // flush grows with its grid and three times over with 10-bit digits, so
  // the two choices interact.
  const int sort_cfgs[][2] = {{8, 15}, {4, 16}, {4, 24}, {8, 8}, {8, 12}, {8, 16}, {16, 8}};
  auto try_cfg = [&](const ocp_tune& t) -> int {
    float ms;
    if (int r = time_cfg(t, &ms)) return r;
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
    return OCP_OK;
  };
  for (int round = 0; round < 2; ++round) {
    const ocp_tune start = best;
    for (const auto& sc : sort_cfgs) {
      ocp_tune t = best;
      t.sort_bits = 8;
      t.sort_waves = sc[0];
      t.sort_kpt = sc[1];
      if (int r = try_cfg(t)) return r;
    }
    for (int msd : {0, 1}) {
      for (int kpt : {8, 16}) {
        ocp_tune t = best;
        t.sort_bits = 10;
        t.sort10_kpt = kpt;
        t.sort_msd = msd;
        if (int r = try_cfg(t)) return r;
      }
    }
    for (int ipt : {8, 16, 32}) {
      ocp_tune t = best;
      t.scan_ipt = ipt;
      if (int r = try_cfg(t)) return r;
    }
    for (int ppt : {4, 8, 16, 32}) {
      for (int grid : {256, 512, 1024, 2048}) {
        ocp_tune t = best;
        t.morton_ppt = ppt;
        t.morton_grid = grid;
        if (int r = try_cfg(t)) return r;
      }
    }
    for (int order : {0, 1}) {
      ocp_tune t = best;
      t.oct_order = order;
      if (int r = try_cfg(t)) return r;
    }
    if (std::memcmp(&start, &best, sizeof(best)) == 0) break;
  }
  if (int r = ocp_set_tune(c, &best)) return r;
  if (chosen) *chosen = best;
  return OCP_OK;
}

// The workload class a tuning result is kept under: device (architecture,
// compute units), plain or batched, point layout, and the point and segment
// counts rounded up to powers of two.
static std::string tune_class(ocp_ctx c, int float3, int64_t n, int32_t n_seg) {
  hipDeviceProp_t prop{};
  std::string arch = "unknown";
  int cus = 0;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
    arch = prop.gcnArchName;
    cus = prop.multiProcessorCount;
  }
  const std::string::size_type colon = arch.find(':');
  if (colon != std::string::npos) arch.resize(colon);
  auto log2_up = [](int64_t v) {
    int k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
  };
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%s/cu%d/%s/f%d/n2^%d/s2^%d", arch.c_str(), cus,
                n_seg == 0 ? "plain" : "batched", float3 ? 3 : 4, log2_up(std::max<int64_t>(n, 1)),
                log2_up(int64_t(std::max(n_seg, 1))));
  return buf;
}

int ocp_autotune_cached(ocp_ctx c, const char* cache_path, const float* points, int float3,
                        int64_t n, const int32_t* offsets, const float* smin, const float* srg,
                        int32_t n_seg, int reps, ocp_tune* chosen, int* from_cache) {
  if (check_ctx(c) || !cache_path) return OCP_ERR_INVALID;
  if (from_cache) *from_cache = 0;
  int64_t total = n;
  if (n_seg > 0) {
    if (!offsets) return fail(c, OCP_ERR_INVALID, "null segment table");
    total = offsets[n_seg];
  }
  const std::string key = tune_class(c, float3, total, n_seg);
  if (FILE* f = std::fopen(cache_path, "r")) {
    char line[512];
    ocp_tune t{}, found{};
    bool hit = false;
    while (std::fgets(line, sizeof(line), f)) {
      char k[256];
      t.oct_order = 0;  // lines written before the octree order was tuned
      t.sort_msd = 0;   // ... or the MSD sort
      const int got = std::sscanf(line, "%255s %d %d %d %d %d %d %d %d %d", k, &t.sort_waves,
                                  &t.sort_kpt, &t.morton_ppt, &t.morton_grid, &t.scan_ipt,
                                  &t.sort_bits, &t.sort10_kpt, &t.oct_order, &t.sort_msd);
      if (got >= 8 && key == k) {
        found = t;  // the last entry for the class wins
        hit = true;
      }
    }
    std::fclose(f);
    if (hit && ocp_set_tune(c, &found) == OCP_OK) {
      if (chosen) *chosen = found;
      if (from_cache) *from_cache = 1;
      return OCP_OK;
    }
  }
  ocp_tune best{};
  if (int r = ocp_autotune(c, points, float3, n, offsets, smin, srg, n_seg, reps, &best))
    return r;
  if (chosen) *chosen = best;
  FILE* f = std::fopen(cache_path, "a");
  if (!f) return fail(c, OCP_ERR_INVALID, std::string("cannot append to ") + cache_path);
  std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d\n", key.c_str(), best.sort_waves,
               best.sort_kpt, best.morton_ppt, best.morton_grid, best.scan_ipt, best.sort_bits,
               best.sort10_kpt, best.oct_order, best.sort_msd);
  std::fclose(f);
  return OCP_OK;
}

// ---------------------------------------------------------------- learned sort
static int learn_model_alloc(ocp_ctx c) {
  LearnState& L = c->learn;
  if (L.model) return OCP_OK;
  int r = 0;
  if ((r = dev_alloc(c, &L.model, kCdfParams)) || (r = dev_alloc(c, &L.meta, 4)) ||
      (r = dev_alloc(c, &L.kleaf, kCdfLeaves + 1)) || (r = dev_alloc(c, &L.fb, kLearnBuckets)) ||
      (r = dev_alloc(c, &L.ctab, kLearnCells + 2)) ||
      (r = dev_alloc(c, &L.loss, 1)))
    return r;
  OCP_HIP(c, hipMemsetAsync(L.model, 0, sizeof(float) * kCdfParams, c->stream));
  OCP_HIP(c, hipMemsetAsync(L.meta, 0, sizeof(uint32_t) * 4, c->stream));
  return OCP_OK;
}

static int learn_tables_alloc(ocp_ctx c, int64_t n) {
  LearnState& L = c->learn;
  LearnTables& t = L.tb;
  if (!t.rows) {
    int r = 0;
    if ((r = dev_alloc(c, &t.rows, int64_t(kLearnHistGrid) * kLearnBuckets)) ||
        (r = dev_alloc(c, &t.hist, int64_t(kLearnReduceGroups) * kLearnBuckets)) ||
        (r = dev_alloc(c, &t.bstart, kLearnBuckets + 1)) || (r = dev_alloc(c, &t.hist_a, 1024)) ||
        (r = dev_alloc(c, &t.hist_b, int64_t(kLearnCoarse) * 1024)) ||
        (r = dev_alloc(c, &t.seg_start, kLearnCoarse + 1)) ||
        (r = dev_alloc(c, &t.tile_start, kLearnCoarse + 1)) || (r = dev_alloc(c, &t.tiles_b, 1)) ||
        (r = dev_alloc(c, &t.info, 8)) || (r = dev_alloc(c, &t.cursor, kLearnCursorWords)))
      return r;
    // The second scatter reads 256 counts per segment and only 64 are ever
    // written; the rest stay zero.
    launch_fill_u32(t.hist_a, 0u, 1024, c->stream);
    launch_fill_u32(t.hist_b, 0u, int64_t(kLearnCoarse) * 1024, c->stream);
    OCP_HIP(c, hipHostMalloc(reinterpret_cast<void**>(&L.h_info), sizeof(int32_t) * 8,
                             hipHostMallocDefault));
    OCP_HIP(c, hipEventCreateWithFlags(&L.ev_info, hipEventDisableTiming));
  }
  if (n > L.n_cap) {
    const int64_t cap = std::max<int64_t>(n, 1 << 20);
    const int32_t big_cap = static_cast<int32_t>(cap / (kLearnBigCap + 1) + 1);
    dev_free(t.big_start); dev_free(t.big_end); dev_free(t.big_tiles); dev_free(t.big_hist);
    dev_free(t.big_cursor);
    int r = 0;
    if ((r = dev_alloc(c, &t.big_start, big_cap)) || (r = dev_alloc(c, &t.big_end, big_cap)) ||
        (r = dev_alloc(c, &t.big_tiles, int64_t(big_cap) + 2)) ||
        (r = dev_alloc(c, &t.big_hist, int64_t(big_cap) * 1024)) ||
        (r = dev_alloc(c, &t.big_cursor, int64_t(big_cap) * 256)))
      return r;
    t.big_cap = big_cap;
    L.n_cap = cap;
  }
  t.fb = L.fb;
  t.ctab = L.ctab;
  t.meta = L.meta;
  return OCP_OK;
}

// Compile the current parameters: leaf starts from the root, then the
// bucket boundaries.
static int learn_compile(ocp_ctx c, bool fit_from_samples, bool linear_leaves) {
  LearnState& L = c->learn;
  launch_lleaf_starts(L.model, L.meta, L.kleaf, c->stream);
  if (fit_from_samples)
    launch_lfit_leaves(L.samp_a, static_cast<int32_t>(L.ring_fill), L.kleaf, L.model, c->stream);
  if (linear_leaves) launch_llinear_leaves(L.kleaf, L.meta, L.model, c->stream);
  launch_lboundaries(L.model, L.meta, L.kleaf, L.fb, L.ctab, c->stream);
  OCP_CHECK_LAUNCH(c);
  L.compiled = true;
  return OCP_OK;
}

int ocp_cdf_reset(ocp_ctx c, uint32_t kmin, uint32_t kmax) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (kmax < kmin) return fail(c, OCP_ERR_INVALID, "kmax < kmin");
  if (int r = learn_model_alloc(c)) return r;
  const uint32_t meta[4] = {kmin, kmax, 0u, 0u};
  OCP_HIP(c, hipMemcpyAsync(c->learn.meta, meta, sizeof(meta), hipMemcpyHostToDevice, c->stream));
  launch_llinear(c->learn.model, c->stream);
  if (int r = learn_compile(c, false, true)) return r;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}

int ocp_cdf_add_samples(ocp_ctx c, const uint32_t* keys, int64_t n, int32_t stride) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n < 0 || stride < 1) return fail(c, OCP_ERR_INVALID, "n < 0 or stride < 1");
  if (n > 0 && !keys) return fail(c, OCP_ERR_INVALID, "null pointer");
  LearnState& L = c->learn;
  if (!L.ring) {
    int r = 0;
    if ((r = dev_alloc(c, &L.ring, kSampleCap)) || (r = dev_alloc(c, &L.samp_a, kSampleCap)) ||
        (r = dev_alloc(c, &L.samp_b, kSampleCap)))
      return r;
  }
  int64_t count = (n + stride - 1) / stride;
  const uint32_t* src = keys;
  if (count > kSampleCap) {  // only the last kSampleCap samples would survive
    src = keys + (count - kSampleCap) * static_cast<int64_t>(stride);
    count = kSampleCap;
  }
  launch_lsample(src, n, stride, L.ring, kSampleCap, L.ring_pos, count, c->stream);
  OCP_CHECK_LAUNCH(c);
  L.ring_pos = (L.ring_pos + count) % kSampleCap;
  L.ring_fill = std::min<int64_t>(kSampleCap, L.ring_fill + count);
  return OCP_OK;
}

int ocp_cdf_clear_samples(ocp_ctx c) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  c->learn.ring_fill = 0;
  c->learn.ring_pos = 0;
  return OCP_OK;
}

int ocp_cdf_sample_count(ocp_ctx c, int64_t* count) {
  if (check_ctx(c) || !count) return OCP_ERR_INVALID;
  *count = c->learn.ring_fill;
  return OCP_OK;
}

int ocp_cdf_train(ocp_ctx c, int32_t steps, float lr, float momentum, uint32_t seed,
                  float* loss_out) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  LearnState& L = c->learn;
  if (L.ring_fill < 2) return fail(c, OCP_ERR_STATE, "fewer than 2 key samples: add samples first");
  if (steps < 0) return fail(c, OCP_ERR_INVALID, "steps < 0");
  if (int r = learn_model_alloc(c)) return r;
  const int32_t ns = static_cast<int32_t>(L.ring_fill);
  OCP_HIP(c, hipMemcpyAsync(L.samp_a, L.ring, sizeof(uint32_t) * ns, hipMemcpyDeviceToDevice,
                            c->stream));
  if (int r = ocp_k_SortKeys(c, L.samp_a, L.samp_b, static_cast<uint32_t>(ns))) return r;
  if (int r = sync_and_check(c)) return r;  // the sorted samples must be valid
  launch_lmeta(L.samp_a, ns, L.meta, c->stream);
  launch_llinear(L.model, c->stream);
  launch_ltrain_root(L.samp_a, ns, L.meta, steps, lr, momentum, seed, L.model, L.loss, c->stream);
  if (int r = learn_compile(c, true, false)) return r;
  float loss = 0.0f;
  OCP_HIP(c, hipMemcpyAsync(&loss, L.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  if (loss_out) *loss_out = steps > 0 ? loss : 0.0f;
  return OCP_OK;
}

int ocp_cdf_get(ocp_ctx c, float* params, uint32_t* kmin, uint32_t* kmax) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  LearnState& L = c->learn;
  if (!L.compiled) return fail(c, OCP_ERR_STATE, "no CDF model: train, set, reset or load one");
  uint32_t meta[4];
  OCP_HIP(c, hipMemcpyAsync(meta, L.meta, sizeof(meta), hipMemcpyDeviceToHost, c->stream));
  if (params)
    OCP_HIP(c, hipMemcpyAsync(params, L.model, sizeof(float) * kCdfParams, hipMemcpyDeviceToHost,
                              c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  if (kmin) *kmin = meta[0];
  if (kmax) *kmax = meta[1];
  return OCP_OK;
}

int ocp_cdf_set(ocp_ctx c, const float* params, uint32_t kmin, uint32_t kmax) {
  if (check_ctx(c) || !params) return OCP_ERR_INVALID;
  if (kmax < kmin) return fail(c, OCP_ERR_INVALID, "kmax < kmin");
  if (int r = learn_model_alloc(c)) return r;
  LearnState& L = c->learn;
  const uint32_t meta[4] = {kmin, kmax, 0u, 0u};
  OCP_HIP(c, hipMemcpyAsync(L.meta, meta, sizeof(meta), hipMemcpyHostToDevice, c->stream));
  OCP_HIP(c, hipMemcpyAsync(L.model, params, sizeof(float) * kCdfParams, hipMemcpyHostToDevice,
                            c->stream));
  if (int r = learn_compile(c, false, false)) return r;
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}

// File: 32-byte header (magic "OCPCDF1", u32 version 1, u32 parameter
// count, u32 kmin, u32 kmax, u32 hidden units, u32 leaves), then the
// parameters as little-endian FP32.
static const char kCdfMagic[8] = {'O', 'C', 'P', 'C', 'D', 'F', '1', '\0'};

int ocp_cdf_save(ocp_ctx c, const char* path) {
  if (check_ctx(c) || !path) return OCP_ERR_INVALID;
  std::vector<float> p(kCdfParams);
  uint32_t kmin = 0, kmax = 0;
  if (int r = ocp_cdf_get(c, p.data(), &kmin, &kmax)) return r;
  const uint32_t hdr[6] = {1u, static_cast<uint32_t>(kCdfParams), kmin, kmax,
                           static_cast<uint32_t>(kCdfHidden), static_cast<uint32_t>(kCdfLeaves)};
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, OCP_ERR_INVALID, std::string("cannot open ") + path);
  const bool ok = std::fwrite(kCdfMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 6, f) == 6 &&
                  std::fwrite(p.data(), 4, p.size(), f) == p.size();
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return fail(c, OCP_ERR_INVALID, std::string("write failed: ") + path);
  return OCP_OK;
}

int ocp_cdf_load(ocp_ctx c, const char* path) {
  if (check_ctx(c) || !path) return OCP_ERR_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, OCP_ERR_INVALID, std::string("cannot open ") + path);
  char magic[8];
  uint32_t hdr[6];
  std::vector<float> p(kCdfParams);
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::fread(hdr, 4, 6, f) == 6;
  ok = ok && std::memcmp(magic, kCdfMagic, 8) == 0 && hdr[0] == 1u &&
       hdr[1] == static_cast<uint32_t>(kCdfParams) && hdr[4] == static_cast<uint32_t>(kCdfHidden) &&
       hdr[5] == static_cast<uint32_t>(kCdfLeaves) && hdr[3] >= hdr[2];
  ok = ok && std::fread(p.data(), 4, p.size(), f) == p.size();
  char extra;
  ok = ok && std::fread(&extra, 1, 1, f) == 0;
  std::fclose(f);
  if (!ok) return fail(c, OCP_ERR_INVALID, std::string("not a CDF model file: ") + path);
  return ocp_cdf_set(c, p.data(), hdr[2], hdr[3]);
}

int ocp_cdf_grad(ocp_ctx c, const float* xs, const float* ts, int32_t n, float* grad) {
  if (check_ctx(c) || !grad) return OCP_ERR_INVALID;
  if (n < 1 || !xs || !ts) return fail(c, OCP_ERR_INVALID, "need n >= 1 samples");
  LearnState& L = c->learn;
  if (!L.model) return fail(c, OCP_ERR_STATE, "no CDF model");
  float* d_grad = nullptr;
  if (int r = dev_alloc(c, &d_grad, kCdfRootParams)) return r;
  launch_fill_u32(reinterpret_cast<uint32_t*>(d_grad), 0u, kCdfRootParams, c->stream);
  launch_lgrad(xs, ts, n, L.model, d_grad, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(grad, d_grad, sizeof(float) * kCdfRootParams, hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_free(d_grad);
  if (e != hipSuccess) return fail(c, OCP_ERR_HIP, hipGetErrorString(e));
  return OCP_OK;
}

int ocp_cdf_buckets(ocp_ctx c, const uint32_t* keys, uint32_t n, int mode, uint32_t* q) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  if (n > 0 && (!keys || !q)) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (mode < 0 || mode > 2) return fail(c, OCP_ERR_INVALID, "mode must be 0, 1 or 2");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  LearnState& L = c->learn;
  if (!L.compiled) return fail(c, OCP_ERR_STATE, "no CDF model: train, set, reset or load one");
  launch_lbucket(keys, static_cast<int32_t>(n), L.model, L.meta, L.kleaf, L.fb, L.ctab, mode, q,
                 c->stream);
  OCP_CHECK_LAUNCH(c);
  return OCP_OK;
}

int ocp_cdf_boundaries(ocp_ctx c, uint32_t* fb) {
  if (check_ctx(c) || !fb) return OCP_ERR_INVALID;
  LearnState& L = c->learn;
  if (!L.compiled) return fail(c, OCP_ERR_STATE, "no CDF model: train, set, reset or load one");
  OCP_HIP(c, hipMemcpyAsync(fb, L.fb, sizeof(uint32_t) * kLearnBuckets, hipMemcpyDeviceToHost,
                            c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  return OCP_OK;
}

int ocp_sort_learned(ocp_ctx c, uint32_t* keys, uint32_t* alt, uint32_t n,
                     ocp_learned_stats* st) {
  if (check_ctx(c)) return OCP_ERR_INVALID;
  ocp_learned_stats local{};
  local.n = n;
  if (n > 0 && (!keys || !alt)) return fail(c, OCP_ERR_INVALID, "null pointer");
  if (n > kMaxPoints) return fail(c, OCP_ERR_CAPACITY, "n > 2^30-1");
  LearnState& L = c->learn;
  if (!L.compiled) return fail(c, OCP_ERR_STATE, "no CDF model: train, set, reset or load one");
  if (n <= 1) {
    if (st) *st = local;
    return OCP_OK;
  }
  if (int r = learn_tables_alloc(c, n)) return r;
  LearnTables& tb = L.tb;
  const int32_t ni = static_cast<int32_t>(n);
  const int stile = sort_tile(c->tune);
  const int64_t fb_tiles = cdiv(n, stile) + tb.big_cap;
  if (int r = ensure_scratch(c, std::max<int64_t>(learn_pass_b_max_tiles(ni), fb_tiles) * 256))
    return r;
  launch_begin(c->ctrl, tb.cursor, kLearnCursorWords, c->stream);
  int32_t rows = 0;
  launch_lhist(keys, n, tb, &rows, c->stream);
  launch_lprefix(tb, rows, stile, c->stream);
  OCP_HIP(c, hipMemcpyAsync(L.h_info, tb.info, sizeof(int32_t) * 8, hipMemcpyDeviceToHost,
                            c->stream));
  OCP_HIP(c, hipEventRecord(L.ev_info, c->stream));
  launch_lpass_a(keys, alt, ni, tb, c->s_flags, c->ctrl, c->stream);
  launch_lpass_b(alt, keys, ni, tb, c->s_flags, c->ctrl, c->stream);
  launch_lsort(keys, tb, c->stream);
  OCP_CHECK_LAUNCH(c);
  // The bucket statistics were copied right after the tables were built;
  // waiting for them does not wait for the scatter passes.
  OCP_HIP(c, hipEventSynchronize(L.ev_info));
  const int32_t n_big = L.h_info[kInfoBig];
  const int32_t big_tiles = L.h_info[kInfoBigTiles];
  local.n_buckets_nonempty = L.h_info[kInfoNonEmpty];
  local.max_bucket = L.h_info[kInfoMaxBucket];
  local.n_big = n_big;
  local.keys_big = static_cast<uint32_t>(L.h_info[kInfoBigKeys]);
  if (n_big > tb.big_cap) return fail(c, OCP_ERR_STATE, "oversized-bucket table overflow");
  // Oversized buckets: segmented OneSweep, in place. Keys reach bucket q
  // only from its range [fb[q], fb[q+1]), so the digits above that range's
  // highest varying bit are constant in the bucket and their passes are
  // identities: buckets whose range varies in the low two digits only
  // (group A) take two passes, the others (group B) four; both counts are
  // even, so the keys end in `keys`. Group B's passes reuse the pass stage
  // ids, so they run in a new epoch (flags of group A's passes then carry a
  // stale tag).
  const int32_t n_a = std::min(L.h_info[kInfoBigA], n_big);
  const int32_t tiles_a = std::clamp(L.h_info[kInfoTilesA], 0, std::max(big_tiles, 0));
  auto fallback = [&](int32_t seg0, int32_t n_seg, int32_t* tile_map, int32_t tiles,
                      int passes, bool new_epoch) {
    if (n_seg <= 0) return;
    if (new_epoch) launch_begin(c->ctrl, nullptr, 0, c->stream);
    uint32_t* hist = tb.big_hist + int64_t(seg0) * 1024;
    uint32_t* cur = tb.big_cursor + int64_t(seg0) * 256;
    // (the histogram kernel zeroes the cursors)
    launch_fill_u32(hist, 0u, int64_t(n_seg) * 1024, c->stream);
    SegView sv{};
    sv.n_seg = n_seg;
    sv.start = tb.big_start + seg0;
    sv.end = tb.big_end + seg0;
    sv.n1 = ni;
    launch_histogram_tiles(keys, sv, tile_map, tiles, stile, hist, c->tune, c->stream, cur);
    uint32_t* src = keys;
    uint32_t* dst = alt;
    for (int pass = 0; pass < passes; ++pass) {
      // (the first pass may reorder equal digits: a cursor scatter)
      if (pass == 0)
        launch_cursor_pass0(src, dst, sv, tile_map, nullptr, tiles, hist, cur, c->ctrl, c->tune,
                            c->stream);
      else
        launch_onesweep_pass(src, dst, sv, tile_map, tiles, hist, c->s_flags, c->ctrl, pass,
                             true, c->tune, c->stream);
      std::swap(src, dst);
    }
    local.fallback_passes += passes;
  };
  fallback(0, n_a, tb.big_tiles, tiles_a, 2, false);
  fallback(n_a, n_big - n_a, tb.big_tiles + n_a + 1, big_tiles - tiles_a, 4, n_a > 0);
  launch_lverify(keys, tb, ni, c->ctrl, c->stream);
  OCP_CHECK_LAUNCH(c);
  uint32_t cw[kCtrlWords];
  OCP_HIP(c, hipMemcpyAsync(cw, c->ctrl, sizeof(cw), hipMemcpyDeviceToHost, c->stream));
  OCP_HIP(c, hipStreamSynchronize(c->stream));
  if (int r = report_timeout(c, cw)) return r;
  local.misordered = static_cast<int32_t>(cw[kCtrlLearnBad]);
  if (local.misordered > 0) {
    if (int r = ocp_k_SortKeys(c, keys, alt, n)) return r;
    OCP_HIP(c, hipStreamSynchronize(c->stream));
    local.full_resort = 1;
  }
  if (st) *st = local;
  return OCP_OK;
}

// ---------------------------------------------------------------- streaming
// Frames rotate over `depth` contexts, each with its own stream, workspace,
// pinned host staging buffer and device points buffer. A frame's upload and
// kernels run in order on its slot's stream, so slots overlap: one frame's
// upload (a DMA engine) and Morton, radix-tree and octree kernels run beside
// another's. Lookback launches (sort passes, unique, edge scan) of two frames
// never overlap: each build's first one waits on the previous frame's last.
struct ocp_pipeline {
  int device = 0;
  int64_t max_points = 0;
  std::vector<ocp_ctx> slots;
  std::vector<hipEvent_t> scan_done;    // after a slot's last lookback launch
  std::vector<hipEvent_t> upload_done;  // after the upload from a slot's host buffer
  std::vector<float*> host;             // pinned, max_points float4 per slot
  std::vector<float*> dev;              // device, max_points float4 per slot
  int32_t last_slot = -1;
  std::string err;
};

static int ppl_fail(ocp_ppl p, int code, const std::string& msg) {
  p->err = msg;
  return code;
}

static int ppl_slot(ocp_ppl p, int32_t slot) {
  if (!p) return OCP_ERR_INVALID;
  if (slot < 0 || slot >= static_cast<int32_t>(p->slots.size()))
    return ppl_fail(p, OCP_ERR_INVALID, "slot out of range");
  return OCP_OK;
}

int ocp_ppl_destroy(ocp_ppl p) {
  if (!p) return OCP_ERR_INVALID;
  for (ocp_ctx c : p->slots)
    if (c) (void)hipStreamSynchronize(c->stream);
  for (float* h : p->host)
    if (h) (void)hipHostFree(h);
  for (float* d : p->dev)
    if (d) (void)hipFree(d);
  for (hipEvent_t e : p->scan_done)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->upload_done)
    if (e) (void)hipEventDestroy(e);
  for (ocp_ctx c : p->slots)
    if (c) (void)ocp_destroy(c);
  delete p;
  return OCP_OK;
}

int ocp_ppl_create(ocp_ppl* out, int device, int32_t depth, int64_t max_points) {
  if (!out) return OCP_ERR_INVALID;
  *out = nullptr;
  if (depth < 1 || depth > 8 || max_points < 0 || max_points > kMaxSegPoints)
    return OCP_ERR_INVALID;
  auto* p = new ocp_pipeline();
  p->device = device;
  p->max_points = max_points;
  p->slots.assign(depth, nullptr);
  p->scan_done.assign(depth, nullptr);
  p->upload_done.assign(depth, nullptr);
  p->host.assign(depth, nullptr);
  p->dev.assign(depth, nullptr);
  const size_t bytes = sizeof(float) * 4 * static_cast<size_t>(max_points);
  for (int32_t i = 0; i < depth; ++i) {
    if (int r = ocp_create(&p->slots[i], device, nullptr)) {
      ocp_ppl_destroy(p);
      return r;
    }
    if (hipEventCreateWithFlags(&p->scan_done[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->upload_done[i], hipEventDisableTiming) != hipSuccess ||
        (bytes && (hipHostMalloc(reinterpret_cast<void**>(&p->host[i]), bytes,
                                 hipHostMallocDefault) != hipSuccess ||
                   hipMalloc(reinterpret_cast<void**>(&p->dev[i]), bytes) != hipSuccess))) {
      ocp_ppl_destroy(p);
      return OCP_ERR_HIP;
    }
  }
  *out = p;
  return OCP_OK;
}

const char* ocp_ppl_last_error(ocp_ppl p) { return p ? p->err.c_str() : "null pipeline"; }

int ocp_ppl_host_buffer(ocp_ppl p, int32_t slot, float** host) {
  if (int r = ppl_slot(p, slot)) return r;
  if (!host) return ppl_fail(p, OCP_ERR_INVALID, "null output");
  if (!p->host[slot]) return ppl_fail(p, OCP_ERR_STATE, "created with max_points = 0");
  if (hipEventSynchronize(p->upload_done[slot]) != hipSuccess)
    return ppl_fail(p, OCP_ERR_HIP, "waiting for the slot's last upload failed");
  *host = p->host[slot];
  return OCP_OK;
}

int ocp_ppl_submit(ocp_ppl p, int32_t slot, const float* points, int float3, int64_t n,
                   float min_coord, float range) {
  if (int r = ppl_slot(p, slot)) return r;
  ocp_ctx c = p->slots[slot];
  const float* pts = points;
  if (!points) {
    if (n < 0 || n > p->max_points)
      return ppl_fail(p, OCP_ERR_CAPACITY, "a staged frame holds at most max_points points");
    const size_t bytes = static_cast<size_t>(n) * (float3 ? 12u : 16u);
    if (bytes) {
      hipError_t e = hipMemcpyAsync(p->dev[slot], p->host[slot], bytes, hipMemcpyHostToDevice,
                                    c->stream);
      if (e == hipSuccess) e = hipEventRecord(p->upload_done[slot], c->stream);
      if (e != hipSuccess)
        return ppl_fail(p, OCP_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    pts = p->dev[slot];
  }
  c->graph_mode = false;
  c->order_wait = p->last_slot >= 0 ? p->scan_done[p->last_slot] : nullptr;
  c->order_record = p->scan_done[slot];
  const int r = ocp_build(c, pts, float3, n, min_coord, range);
  c->order_wait = nullptr;
  c->order_record = nullptr;
  if (r) return ppl_fail(p, r, c->err);
  p->last_slot = slot;
  return OCP_OK;
}

int ocp_ppl_wait(ocp_ppl p, int32_t slot) {
  if (int r = ppl_slot(p, slot)) return r;
  if (int r = ocp_synchronize(p->slots[slot])) return ppl_fail(p, r, p->slots[slot]->err);
  return OCP_OK;
}

int ocp_ppl_context(ocp_ppl p, int32_t slot, ocp_ctx* ctx) {
  if (int r = ppl_slot(p, slot)) return r;
  if (!ctx) return ppl_fail(p, OCP_ERR_INVALID, "null output");
  *ctx = p->slots[slot];
  return OCP_OK;
}

}  // extern "C"

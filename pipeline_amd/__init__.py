"""MI355X (gfx950) octree-construction pipeline - Python host binding.

A thin ctypes layer over the C ABI in ``include/octree_pipeline.h``
(``pipeline_amd/_lib/libocpipe.so``). It mirrors the reference's per-stage
entry points (``k_ComputeMortonCode``, ``k_SortKeys``/``k_SimpleRadixSort``,
``k_CountUnique``, ``k_BuildRadixTree``, ``k_EdgeCount``, ``k_PartialSum``,
``k_MakeOctNodes``, ``k_LinkLeafNodes``; include/kernels/all.hpp and
include/cuda/kernels/all.cuh of ucsc-redwood/pipeline) with the same
argument meaning, and exposes the fused device-resident build.

There is no CPU fallback: importing works without a GPU, but creating a
``Context`` loads the HIP library and fails loudly when it is missing or no
device is present.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libocpipe.so")

OCT_NODE_DTYPE = np.dtype(
    [
        ("children", "<i4", (8,)),
        ("corner", "<f4", (4,)),
        ("cell_size", "<f4"),
        ("child_node_mask", "<i4"),
        ("child_leaf_mask", "<i4"),
    ]
)
assert OCT_NODE_DTYPE.itemsize == 60

OCP_OK = 0

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_F = ctypes.c_float
_INT = ctypes.c_int
_SZ = ctypes.c_size_t


class OcpError(RuntimeError):
    pass


class Results(ctypes.Structure):
    _fields_ = [
        ("sorted_keys", _P), ("unique_keys", _P), ("prefix_n", _P),
        ("has_leaf_left", _P), ("has_leaf_right", _P), ("left_child", _P),
        ("parent", _P), ("edge_count", _P), ("offsets", _P), ("nodes", _P),
        ("n_unique", _P), ("n_oct", _P), ("oct_start", _P),
        ("n_seg", _I32), ("n_points", _I64), ("seg_min", _P), ("seg_range", _P),
        ("sorted_index", _P), ("unique_index", _P),
    ]


class Tune(ctypes.Structure):
    _fields_ = [("sort_waves", _I32), ("sort_kpt", _I32), ("morton_ppt", _I32),
                ("morton_grid", _I32), ("scan_ipt", _I32), ("sort_bits", _I32),
                ("sort10_kpt", _I32), ("oct_order", _I32), ("sort_msd", _I32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class LearnedStats(ctypes.Structure):
    _fields_ = [("n", _I64), ("n_buckets_nonempty", _I32), ("max_bucket", _I32),
                ("n_big", _I32), ("misordered", _I32), ("keys_big", _I64),
                ("full_resort", _I32), ("fallback_passes", _I32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class WorkspaceSizes(ctypes.Structure):
    _fields_ = [("build_bytes", _I64), ("sort_bytes", _I64), ("unique_bytes", _I64),
                ("scan_bytes", _I64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class TreeReport(ctypes.Structure):
    """ocp_tree_report: counted violations of the octree's structure, plus
    the node-mask links and those whose child lies in the parent's octant."""
    _fields_ = [(f, _I64) for f in (
        "slots", "written", "covered_unwritten", "written_uncovered", "chain_past_end",
        "bad_w", "bad_cell", "off_grid", "leaf_mask", "bad_root", "links",
        "links_in_octant")]
    VIOLATIONS = ("covered_unwritten", "written_uncovered", "chain_past_end", "bad_w",
                  "bad_cell", "off_grid", "leaf_mask", "bad_root")

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# ocp_compact_node (12 bytes).
COMPACT_NODE_DTYPE = np.dtype([("prefix", np.uint32), ("level", np.uint8),
                               ("node_mask", np.uint8), ("child_set", np.uint8),
                               ("leaf_mask", np.uint8), ("first_child", np.uint32)])
assert COMPACT_NODE_DTYPE.itemsize == 12

STATIC_SORT_MAX_KEYS = 1 << 24
CDF_PARAMS = 12338
CDF_ROOT_PARAMS = 49
CDF_BUCKETS = 16384
CDF_SAMPLE_CAP = 1 << 22


# Every function the C ABI exports, with its ctypes signature. The CPU test
# suite checks that the built library exports exactly these.
SIGNATURES = {
    "ocp_create": (_INT, [ctypes.POINTER(_P), _INT, _P]),
    "ocp_destroy": (_INT, [_P]),
    "ocp_last_error": (ctypes.c_char_p, [_P]),
    "ocp_synchronize": (_INT, [_P]),
    "ocp_set_wait_limit": (_INT, [_P, _I64]),
    "ocp_malloc": (_INT, [_P, ctypes.POINTER(_P), _SZ]),
    "ocp_device_memory": (_INT, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "ocp_free": (_INT, [_P, _P]),
    "ocp_memcpy_h2d": (_INT, [_P, _P, _P, _SZ]),
    "ocp_memcpy_d2h": (_INT, [_P, _P, _P, _SZ]),
    "ocp_memset": (_INT, [_P, _P, _INT, _SZ]),
    "ocp_copy": (_INT, [_P, _P, _P, _SZ]),
    "ocp_k_InitRandomVec4": (_INT, [_P, _P, _I32, _F, _F, _I32]),
    "ocp_k_ComputeMortonCode": (_INT, [_P, _P, _P, _SZ, _F, _F]),
    "ocp_k_ComputeMortonCode3": (_INT, [_P, _P, _P, _SZ, _F, _F]),
    "ocp_k_GlobalHistogram": (_INT, [_P, _P, _P, _U32]),
    "ocp_k_DigitBinning": (_INT, [_P, _P, _P, _P, _U32, _U32]),
    "ocp_k_SortKeys": (_INT, [_P, _P, _P, _U32]),
    "ocp_k_SortPairs": (_INT, [_P, _P, _P, _P, _P, _U32]),
    "ocp_k_CountUniquePairs": (_INT, [_P, _P, _P, _P, _I32]),
    "ocp_set_permutation": (_INT, [_P, _INT]),
    "ocp_k_SortKeys_static": (_INT, [_P, _P, _P, _U32]),
    "ocp_k_CountUnique": (_INT, [_P, _P, _P, _I32]),
    "ocp_k_BuildRadixTree": (_INT, [_P, _I32, _P, _P, _P, _P, _P, _P]),
    "ocp_k_EdgeCount": (_INT, [_P, _P, _P, _P, _I32]),
    "ocp_k_PartialSum": (_INT, [_P, _P, _I64, _I64, _P]),
    "ocp_k_InitOctNodes": (_INT, [_P, _P, _I32, _P, _P, _I32, _F, _F]),
    "ocp_k_MakeOctNodes": (_INT, [_P, _P, _P, _P, _P, _P, _P, _F, _F, _I32, _I32]),
    "ocp_k_LinkLeafNodes": (_INT, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32]),
    "ocp_build": (_INT, [_P, _P, _INT, _I64, _F, _F]),
    "ocp_build_batched": (_INT, [_P, _P, _INT, _P, _P, _P, _I32]),
    "ocp_build_autobox": (_INT, [_P, _P, _INT, _I64]),
    "ocp_transform_points": (_INT, [_P, _P, _P, _INT, _I64, _P, _F, _F]),
    "ocp_check_batch": (_INT, [_P, _P, _P, _I32]),
    "ocp_workspace_size": (_INT, [_P, _I64, _I32, _INT, ctypes.POINTER(WorkspaceSizes)]),
    "ocp_set_graph_mode": (_INT, [_P, _INT]),
    "ocp_get_results": (_INT, [_P, ctypes.POINTER(Results)]),
    "ocp_get_counts": (_INT, [_P, _P, _P]),
    "ocp_ppl_create": (_INT, [ctypes.POINTER(_P), _INT, _I32, _I64]),
    "ocp_ppl_destroy": (_INT, [_P]),
    "ocp_ppl_last_error": (ctypes.c_char_p, [_P]),
    "ocp_ppl_host_buffer": (_INT, [_P, _I32, ctypes.POINTER(_P)]),
    "ocp_ppl_submit": (_INT, [_P, _I32, _P, _INT, _I64, _F, _F]),
    "ocp_ppl_wait": (_INT, [_P, _I32]),
    "ocp_ppl_context": (_INT, [_P, _I32, ctypes.POINTER(_P)]),
    "ocp_validate_octree": (_INT, [_P, ctypes.POINTER(TreeReport)]),
    "ocp_compact_octree": (_INT, [_P, _P, _I64, _P, _I64, ctypes.POINTER(_I64),
                                  ctypes.POINTER(_I64)]),
    "ocp_expand_octree": (_INT, [_P, _P, _P, _I64, _P, _I32, _P, _P, _F, _F, _P]),
    "ocp_set_stage_timing": (_INT, [_P, _INT]),
    "ocp_set_cold_timing": (_INT, [_P, _I64]),
    "ocp_get_stage_ms": (_INT, [_P, _P]),
    "ocp_get_kernel_ms": (_INT, [_P, _P, _P, _I32, ctypes.POINTER(_I32)]),
    "ocp_time_builds": (_INT, [_P, _P, _INT, _P, _P, _P, _I32, _INT, _INT, ctypes.POINTER(_F)]),
    "ocp_get_tune": (_INT, [_P, ctypes.POINTER(Tune)]),
    "ocp_set_tune": (_INT, [_P, ctypes.POINTER(Tune)]),
    "ocp_autotune": (_INT, [_P, _P, _INT, _I64, _P, _P, _P, _I32, _INT, ctypes.POINTER(Tune)]),
    "ocp_autotune_cached": (_INT, [_P, ctypes.c_char_p, _P, _INT, _I64, _P, _P, _P, _I32, _INT,
                                   ctypes.POINTER(Tune), ctypes.POINTER(_INT)]),
    "ocp_cdf_reset": (_INT, [_P, _U32, _U32]),
    "ocp_cdf_add_samples": (_INT, [_P, _P, _I64, _I32]),
    "ocp_cdf_clear_samples": (_INT, [_P]),
    "ocp_cdf_sample_count": (_INT, [_P, ctypes.POINTER(_I64)]),
    "ocp_cdf_train": (_INT, [_P, _I32, _F, _F, _U32, ctypes.POINTER(_F)]),
    "ocp_cdf_get": (_INT, [_P, _P, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "ocp_cdf_set": (_INT, [_P, _P, _U32, _U32]),
    "ocp_cdf_save": (_INT, [_P, ctypes.c_char_p]),
    "ocp_cdf_load": (_INT, [_P, ctypes.c_char_p]),
    "ocp_cdf_grad": (_INT, [_P, _P, _P, _I32, _P]),
    "ocp_cdf_buckets": (_INT, [_P, _P, _U32, _INT, _P]),
    "ocp_cdf_boundaries": (_INT, [_P, _P]),
    "ocp_sort_learned": (_INT, [_P, _P, _P, _U32, ctypes.POINTER(LearnedStats)]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libocpipe.so (raises OcpError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OcpError(
            f"HIP library {path} is missing: build it with "
            "`python -m pipeline_amd.build` (there is no CPU fallback)")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def workspace_size(n: int, n_seg: int = 0, permutation: bool = False, tune: dict | None = None):
    """Device bytes a context allocates for a build of n points (and the
    per-stage sort, unique and scan scratch); no GPU used."""
    out = WorkspaceSizes()
    t = None
    if tune is not None:
        t = Tune(**{k: int(v) for k, v in tune.items()})
    rc = load_library().ocp_workspace_size(ctypes.byref(t) if t is not None else None, int(n),
                                           int(n_seg), int(bool(permutation)), ctypes.byref(out))
    if rc != OCP_OK:
        raise OcpError(f"ocp_workspace_size: status {rc}")
    return out.as_dict()


def check_batch(seg_offsets, seg_min, seg_range) -> int:
    """Status the batched build would return for this table (no GPU used)."""
    offs = np.ascontiguousarray(seg_offsets, np.int32)
    mn = np.ascontiguousarray(seg_min, np.float32)
    rg = np.ascontiguousarray(seg_range, np.float32)
    return int(load_library().ocp_check_batch(_ptr(offs), _ptr(mn), _ptr(rg), offs.size - 1))


class DeviceArray:
    """A device allocation with a numpy dtype/shape, owned by a Context."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = ctx._malloc(self.nbytes)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    def copy_from(self, host: np.ndarray) -> "DeviceArray":
        h = np.ascontiguousarray(host, dtype=self.dtype)
        if h.nbytes > self.nbytes:
            raise OcpError("host array larger than the device buffer")
        self.ctx._check(self.ctx.lib.ocp_memcpy_h2d(self.ctx.h, self.ptr, _ptr(h), h.nbytes))
        return self

    def to_host(self, count: int | None = None) -> np.ndarray:
        n = self.size if count is None else int(count)
        out = np.empty(n, self.dtype)
        if n:
            self.ctx._check(self.ctx.lib.ocp_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr,
                                                        n * self.dtype.itemsize))
        return out.reshape(self.shape) if count is None else out

    def zero(self) -> "DeviceArray":
        self.ctx._check(self.ctx.lib.ocp_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            self.ctx.lib.ocp_free(self.ctx.h, self.ptr)
            self.ptr = None


def d2h(ctx: "Context", ptr: int, count: int, dtype) -> np.ndarray:
    """Copy `count` elements of `dtype` from a raw device pointer."""
    dt = np.dtype(dtype)
    out = np.empty(int(count), dt)
    if count:
        ctx._check(ctx.lib.ocp_memcpy_d2h(ctx.h, _ptr(out), ctypes.c_void_p(ptr),
                                          out.nbytes))
    return out


@dataclass
class BuildOutput:
    """Host copies of one build's stage outputs, per segment."""
    sorted_keys: list
    unique_keys: list
    prefix_n: list
    has_leaf_left: list
    has_leaf_right: list
    left_child: list
    parent: list
    edge_count: list
    offsets: list
    nodes: list
    n_unique: np.ndarray
    n_oct: np.ndarray
    sorted_index: list | None = None   # with the permutation on
    unique_index: list | None = None


class Context:
    """One HIP stream + workspace (ocp_ctx). Not thread safe."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.ocp_create(ctypes.byref(h), device, None)
        if rc != OCP_OK:
            raise OcpError(f"ocp_create failed with status {rc} (no usable GPU?)")
        self.h = h
        self._owned: list[DeviceArray] = []

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int):
        if rc != OCP_OK:
            msg = self.lib.ocp_last_error(self.h)
            raise OcpError(f"status {rc}: {msg.decode() if msg else ''}")

    def _malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._check(self.lib.ocp_malloc(self.h, ctypes.byref(p), max(int(nbytes), 1)))
        return p.value

    def empty(self, shape, dtype) -> DeviceArray:
        a = DeviceArray(self, shape, dtype)
        self._owned.append(a)
        return a

    def zeros(self, shape, dtype) -> DeviceArray:
        return self.empty(shape, dtype).zero()

    def upload(self, host: np.ndarray, dtype=None) -> DeviceArray:
        h = np.ascontiguousarray(host, dtype=dtype or host.dtype)
        return self.empty(h.shape, h.dtype).copy_from(h)

    def device_memory(self):
        """(free, total) bytes of the context's device."""
        fr, tot = _I64(), _I64()
        self._check(self.lib.ocp_device_memory(self.h, ctypes.byref(fr), ctypes.byref(tot)))
        return int(fr.value), int(tot.value)

    def synchronize(self):
        self._check(self.lib.ocp_synchronize(self.h))

    def set_wait_limit(self, polls: int):
        """Polls after which a lookback wait gives up (-1: the default)."""
        self._check(self.lib.ocp_set_wait_limit(self.h, int(polls)))

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int):
        """Device-to-device copy kernel (16-byte multiples)."""
        self._check(self.lib.ocp_copy(self.h, dst.ptr, src.ptr, nbytes))

    def close(self):
        if getattr(self, "h", None):
            for a in self._owned:
                a.free()
            self._owned.clear()
            self.lib.ocp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ stages
    def k_InitRandomVec4(self, data: DeviceArray, n: int, min_coord: float, range_: float,
                         seed: int):
        """The reference's GPU generator (kernels/cuda/00_init.cu:27-50)."""
        self._check(self.lib.ocp_k_InitRandomVec4(self.h, data.ptr, int(n), float(min_coord),
                                                  float(range_), int(seed)))

    def k_ComputeMortonCode(self, data: DeviceArray, keys: DeviceArray, n: int,
                            min_coord: float, range_: float):
        fn = (self.lib.ocp_k_ComputeMortonCode3 if data.shape[-1] == 3
              else self.lib.ocp_k_ComputeMortonCode)
        self._check(fn(self.h, data.ptr, keys.ptr, n, min_coord, range_))

    def k_GlobalHistogram(self, keys: DeviceArray, hist: DeviceArray, n: int):
        self._check(self.lib.ocp_k_GlobalHistogram(self.h, keys.ptr, hist.ptr, n))

    def k_DigitBinning(self, hist, keys, alt, n: int, radix_shift: int):
        self._check(self.lib.ocp_k_DigitBinning(self.h, hist.ptr, keys.ptr, alt.ptr, n,
                                                radix_shift))

    def k_SortKeys(self, keys: DeviceArray, keys_alt: DeviceArray, n: int):
        self._check(self.lib.ocp_k_SortKeys(self.h, keys.ptr, keys_alt.ptr, n))

    k_SimpleRadixSort = k_SortKeys

    def k_SortKeys_static(self, keys: DeviceArray, keys_alt: DeviceArray, n: int):
        """ocp_k_SortKeys with its scratch in the library's fixed device array
        (no allocation; at most STATIC_SORT_MAX_KEYS keys)."""
        self._check(self.lib.ocp_k_SortKeys_static(self.h, keys.ptr, keys_alt.ptr, n))

    def k_SortPairs(self, keys, vals, keys_alt, vals_alt, n: int):
        self._check(self.lib.ocp_k_SortPairs(self.h, keys.ptr, vals.ptr, keys_alt.ptr,
                                             vals_alt.ptr, n))

    def k_CountUniquePairs(self, keys, vals, n: int) -> int:
        out = self.empty(1, np.int32)
        self._check(self.lib.ocp_k_CountUniquePairs(self.h, keys.ptr, vals.ptr, out.ptr, n))
        u = int(out.to_host()[0])
        out.free()
        self._owned.remove(out)
        return u

    def set_permutation(self, enable: bool):
        self._check(self.lib.ocp_set_permutation(self.h, int(bool(enable))))

    def k_CountUnique(self, keys: DeviceArray, n: int) -> int:
        out = self.empty(1, np.int32)
        self._check(self.lib.ocp_k_CountUnique(self.h, keys.ptr, out.ptr, n))
        u = int(out.to_host()[0])
        out.free()
        self._owned.remove(out)
        return u

    def k_BuildRadixTree(self, n: int, codes, prefix_n, has_leaf_left,
                         has_leaf_right, left_child, parent):
        self._check(self.lib.ocp_k_BuildRadixTree(
            self.h, n, codes.ptr, prefix_n.ptr, has_leaf_left.ptr, has_leaf_right.ptr,
            left_child.ptr, parent.ptr))

    def k_EdgeCount(self, prefix_n, parents, edge_count, n_brt_nodes: int):
        self._check(self.lib.ocp_k_EdgeCount(self.h, prefix_n.ptr, parents.ptr,
                                             edge_count.ptr, n_brt_nodes))

    def k_PartialSum(self, data, first: int, last: int, d_first):
        self._check(self.lib.ocp_k_PartialSum(self.h, data.ptr, first, last, d_first.ptr))

    def k_InitOctNodes(self, nodes, n_alloc, codes, prefix_n, n_brt, min_coord, range_):
        self._check(self.lib.ocp_k_InitOctNodes(
            self.h, nodes.ptr, n_alloc, codes.ptr, prefix_n.ptr if prefix_n else None,
            n_brt, min_coord, range_))

    def k_MakeOctNodes(self, nodes, node_offsets, rt_node_counts, codes, rt_prefixN,
                       rt_parents, min_coord, range_, N, n_alloc):
        self._check(self.lib.ocp_k_MakeOctNodes(
            self.h, nodes.ptr, node_offsets.ptr, rt_node_counts.ptr, codes.ptr,
            rt_prefixN.ptr, rt_parents.ptr, min_coord, range_, N, n_alloc))

    def k_LinkLeafNodes(self, nodes, node_offsets, rt_node_counts, codes,
                        rt_hasLeafLeft, rt_hasLeafRight, rt_prefixN, rt_parents,
                        rt_leftChild, N, n_alloc):
        self._check(self.lib.ocp_k_LinkLeafNodes(
            self.h, nodes.ptr, node_offsets.ptr, rt_node_counts.ptr, codes.ptr,
            rt_hasLeafLeft.ptr, rt_hasLeafRight.ptr, rt_prefixN.ptr, rt_parents.ptr,
            rt_leftChild.ptr, N, n_alloc))

    # ------------------------------------------------------------ pipeline
    def set_graph_mode(self, enable: bool):
        self._check(self.lib.ocp_set_graph_mode(self.h, int(bool(enable))))

    def set_stage_timing(self, enable: bool):
        self._check(self.lib.ocp_set_stage_timing(self.h, int(bool(enable))))

    def set_cold_timing(self, flush_bytes: int):
        """Overwrite a buffer of flush_bytes before every timed kernel (0: off)."""
        self._check(self.lib.ocp_set_cold_timing(self.h, int(flush_bytes)))

    def stage_ms(self) -> dict:
        ms = np.zeros(7, np.float32)
        self._check(self.lib.ocp_get_stage_ms(self.h, _ptr(ms)))
        names = ["morton", "sort", "unique", "radix_tree", "edge_scan", "octree", "total"]
        return dict(zip(names, ms.astype(float).tolist()))

    def kernel_ms(self) -> dict:
        """Per-kernel ms of the last eager build (stage timing on)."""
        n = ctypes.c_int32()
        ms = np.zeros(64, np.float32)
        names = (ctypes.c_char_p * 64)()
        self._check(self.lib.ocp_get_kernel_ms(self.h, _ptr(ms), names, 64, ctypes.byref(n)))
        return {names[i].decode(): float(ms[i]) for i in range(n.value)}

    def build(self, points: DeviceArray, n: int | None = None, min_coord=0.0,
              range_=1024.0):
        n = points.shape[0] if n is None else n
        f3 = int(points.shape[-1] == 3)
        self._check(self.lib.ocp_build(self.h, points.ptr, f3, n, min_coord, range_))

    def build_batched(self, points: DeviceArray, seg_offsets, seg_min=None, seg_range=None):
        """seg_min/seg_range None: each cloud's cube box is reduced on the device."""
        offs = np.ascontiguousarray(seg_offsets, np.int32)
        f3 = int(points.shape[-1] == 3)
        if seg_min is None and seg_range is None:
            self._check(self.lib.ocp_build_batched(self.h, points.ptr, f3, _ptr(offs), None,
                                                   None, offs.size - 1))
            return
        mn = np.ascontiguousarray(seg_min, np.float32)
        rg = np.ascontiguousarray(seg_range, np.float32)
        self._check(self.lib.ocp_build_batched(self.h, points.ptr, f3, _ptr(offs),
                                               _ptr(mn), _ptr(rg), offs.size - 1))

    def build_autobox(self, points: DeviceArray, n: int | None = None):
        n = points.shape[0] if n is None else n
        f3 = int(points.shape[-1] == 3)
        self._check(self.lib.ocp_build_autobox(self.h, points.ptr, f3, int(n)))

    def transform_points(self, src: DeviceArray, dst: DeviceArray, affine, lo: float,
                         hi: float, n: int | None = None):
        """dst = clamp(A src + t, lo, hi) per point on the device; affine is
        a 3x4 [A | t] (the fourth float of a float4 point passes through)."""
        n = src.shape[0] if n is None else n
        m = np.ascontiguousarray(np.asarray(affine, np.float64).reshape(3, 4), np.float64)
        flat = np.concatenate([m[:, :3].ravel(), m[:, 3]]).astype(np.float32)
        self._check(self.lib.ocp_transform_points(self.h, src.ptr, dst.ptr,
                                                  int(src.shape[-1] == 3), int(n), _ptr(flat),
                                                  float(lo), float(hi)))

    def boxes(self):
        """(seg_min, seg_range) host copies of the boxes the last batched or
        autobox build used."""
        r = self.results()
        if not r.seg_min:
            raise OcpError("the last build took its box from the caller")
        return (d2h(self, r.seg_min, r.n_seg, np.float32), d2h(self, r.seg_range, r.n_seg,
                                                                np.float32))

    def results(self) -> Results:
        r = Results()
        self._check(self.lib.ocp_get_results(self.h, ctypes.byref(r)))
        return r

    def counts(self, n_seg: int):
        u = np.zeros(n_seg, np.int32)
        k = np.zeros(n_seg, np.int32)
        self._check(self.lib.ocp_get_counts(self.h, _ptr(u), _ptr(k)))
        return u, k

    def validate_octree(self) -> dict:
        """Structural check of the last build's nodes on the device (counts
        per kind of violation; see TreeReport.VIOLATIONS)."""
        rep = TreeReport()
        self._check(self.lib.ocp_validate_octree(self.h, ctypes.byref(rep)))
        return rep.as_dict()

    def compact_octree(self):
        """(nodes, children) device arrays of the last build's compact form."""
        kn, kc = _I64(), _I64()
        self._check(self.lib.ocp_compact_octree(self.h, None, 0, None, 0, ctypes.byref(kn),
                                                ctypes.byref(kc)))
        nodes = self.empty(max(kn.value, 1), COMPACT_NODE_DTYPE)
        children = self.empty(max(kc.value, 1), np.int32)
        self._check(self.lib.ocp_compact_octree(self.h, nodes.ptr, kn.value, children.ptr,
                                                kc.value, ctypes.byref(kn), ctypes.byref(kc)))
        nodes.shape, children.shape = (kn.value,), (kc.value,)
        return nodes, children

    def expand_octree(self, nodes: DeviceArray, children: DeviceArray, oct_start: int,
                      n_seg: int, seg_min: int | None = None, seg_range: int | None = None,
                      min_coord: float = 0.0, range_: float = 1024.0) -> DeviceArray:
        """60-byte nodes from a compact form; oct_start, seg_min and seg_range
        are device pointers (ocp_results fields). Returns a device array."""
        out = self.empty(max(nodes.size, 1) * 60, np.uint8)
        self._check(self.lib.ocp_expand_octree(self.h, nodes.ptr, children.ptr, children.size,
                                               ctypes.c_void_p(oct_start), int(n_seg),
                                               ctypes.c_void_p(seg_min) if seg_min else None,
                                               ctypes.c_void_p(seg_range) if seg_range else None,
                                               float(min_coord), float(range_), out.ptr))
        return out

    def fetch(self, seg_offsets=None) -> BuildOutput:
        """Copy every stage output of the last build to the host."""
        r = self.results()
        n_seg = r.n_seg
        offs = (np.array([0, r.n_points], np.int64) if seg_offsets is None
                else np.asarray(seg_offsets, np.int64))
        u, k = self.counts(n_seg)
        total = int(r.n_points)
        sk = d2h(self, r.sorted_keys, total, np.uint32)
        uk = d2h(self, r.unique_keys, total, np.uint32)
        pn = d2h(self, r.prefix_n, total, np.uint8)
        hl = d2h(self, r.has_leaf_left, total, np.uint8)
        hr = d2h(self, r.has_leaf_right, total, np.uint8)
        lc = d2h(self, r.left_child, total, np.int32)
        par = d2h(self, r.parent, total, np.int32)
        e = d2h(self, r.edge_count, total, np.int32)
        p = d2h(self, r.offsets, total, np.int32)
        ost = d2h(self, r.oct_start, n_seg + 1, np.int32)
        nodes = d2h(self, r.nodes, int(ost[-1]) * 60, np.uint8).view(OCT_NODE_DTYPE)
        out = BuildOutput([], [], [], [], [], [], [], [], [], [], u, k)
        si = ui = None
        if r.sorted_index:
            si = d2h(self, r.sorted_index, total, np.uint32)
            ui = d2h(self, r.unique_index, total, np.uint32)
            out.sorted_index, out.unique_index = [], []
        for s in range(n_seg):
            a, b = int(offs[s]), int(offs[s + 1])
            m = max(int(u[s]) - 1, 0)
            out.sorted_keys.append(sk[a:b])
            out.unique_keys.append(uk[a:a + int(u[s])])
            out.prefix_n.append(pn[a:a + m])
            out.has_leaf_left.append(hl[a:a + m])
            out.has_leaf_right.append(hr[a:a + m])
            out.left_child.append(lc[a:a + m])
            out.parent.append(par[a:a + m])
            out.edge_count.append(e[a:a + m])
            out.offsets.append(p[a:a + m])
            out.nodes.append(nodes[int(ost[s]):int(ost[s]) + int(k[s])])
            if si is not None:
                out.sorted_index.append(si[a:b])
                out.unique_index.append(ui[a:a + int(u[s])])
        return out

    def time_builds(self, points: DeviceArray, warmup: int, steps: int,
                    seg_offsets=None, seg_min=None, seg_range=None, n=None,
                    min_coord=0.0, range_=1024.0, first: int = 0) -> float:
        """Mean ms per build over `steps` builds after `warmup` (plain when
        seg_offsets is None: n points from point index `first`)."""
        f3 = int(points.shape[-1] == 3)
        base = points.ptr + first * points.shape[-1] * 4
        ms = ctypes.c_float()
        if seg_offsets is None:
            n = points.shape[0] if n is None else n
            offs = np.array([0, n], np.int32)
            mn = np.array([min_coord], np.float32)
            rg = np.array([range_], np.float32)
            n_seg = 0
        else:
            offs = np.ascontiguousarray(seg_offsets, np.int32)
            mn = np.ascontiguousarray(seg_min, np.float32)
            rg = np.ascontiguousarray(seg_range, np.float32)
            n_seg = offs.size - 1
        self._check(self.lib.ocp_time_builds(self.h, ctypes.c_void_p(base), f3, _ptr(offs),
                                             _ptr(mn), _ptr(rg), n_seg, warmup, steps,
                                             ctypes.byref(ms)))
        return float(ms.value)

    def get_tune(self) -> dict:
        t = Tune()
        self._check(self.lib.ocp_get_tune(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tune(self, **kw):
        t = Tune()
        self._check(self.lib.ocp_get_tune(self.h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(self.lib.ocp_set_tune(self.h, ctypes.byref(t)))

    def autotune(self, points: DeviceArray, n=None, seg_offsets=None, seg_min=None,
                 seg_range=None, reps=3, cache: str | None = None) -> dict:
        """Time the tile configurations and digit widths on this workload and
        keep the fastest. With `cache` (a text file), a workload class already
        tuned there is applied without timing (self.tune_from_cache)."""
        t = Tune()
        f3 = int(points.shape[-1] == 3)
        if seg_offsets is None:
            n = points.shape[0] if n is None else n
            offs = None
            mn = np.array([0.0], np.float32)
            rg = np.array([1024.0], np.float32)
            args = (points.ptr, f3, int(n), None, _ptr(mn), _ptr(rg), 0, reps, ctypes.byref(t))
        else:
            offs = np.ascontiguousarray(seg_offsets, np.int32)
            mn = np.ascontiguousarray(seg_min, np.float32)
            rg = np.ascontiguousarray(seg_range, np.float32)
            args = (points.ptr, f3, int(offs[-1]), _ptr(offs), _ptr(mn), _ptr(rg),
                    offs.size - 1, reps, ctypes.byref(t))
        self.tune_from_cache = False
        if cache is None:
            self._check(self.lib.ocp_autotune(self.h, *args))
        else:
            hit = ctypes.c_int(0)
            self._check(self.lib.ocp_autotune_cached(self.h, str(cache).encode(), *args,
                                                     ctypes.byref(hit)))
            self.tune_from_cache = bool(hit.value)
        return t.as_dict()

    # ---- learned-CDF bucket sort
    def cdf_reset(self, kmin: int = 0, kmax: int = (1 << 30) - 1):
        """Linear CDF over [kmin, kmax] (the untrained model)."""
        self._check(self.lib.ocp_cdf_reset(self.h, int(kmin), int(kmax)))

    def cdf_add_samples(self, keys: DeviceArray, n: int | None = None, stride: int = 100):
        n = keys.size if n is None else n
        self._check(self.lib.ocp_cdf_add_samples(self.h, keys.ptr, int(n), int(stride)))

    def cdf_clear_samples(self):
        self._check(self.lib.ocp_cdf_clear_samples(self.h))

    def cdf_sample_count(self) -> int:
        v = _I64()
        self._check(self.lib.ocp_cdf_sample_count(self.h, ctypes.byref(v)))
        return int(v.value)

    def cdf_train(self, steps: int = 2000, lr: float = 1.0, momentum: float = 0.9,
                  seed: int = 1) -> float:
        loss = _F()
        self._check(self.lib.ocp_cdf_train(self.h, int(steps), float(lr), float(momentum),
                                           int(seed), ctypes.byref(loss)))
        return float(loss.value)

    def cdf_get(self):
        """(params float32[CDF_PARAMS], kmin, kmax)"""
        p = np.empty(CDF_PARAMS, np.float32)
        lo, hi = _U32(), _U32()
        self._check(self.lib.ocp_cdf_get(self.h, _ptr(p), ctypes.byref(lo), ctypes.byref(hi)))
        return p, int(lo.value), int(hi.value)

    def cdf_set(self, params: np.ndarray, kmin: int, kmax: int):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != CDF_PARAMS:
            raise ValueError(f"expected {CDF_PARAMS} parameters, got {p.size}")
        self._check(self.lib.ocp_cdf_set(self.h, _ptr(p), int(kmin), int(kmax)))

    def cdf_save(self, path: str):
        self._check(self.lib.ocp_cdf_save(self.h, os.fsencode(path)))

    def cdf_load(self, path: str):
        self._check(self.lib.ocp_cdf_load(self.h, os.fsencode(path)))

    def cdf_grad(self, xs: np.ndarray, ts: np.ndarray) -> np.ndarray:
        """Mean gradient of 0.5 (r(x) - t)^2 over the current root parameters."""
        dx = self.upload(np.ascontiguousarray(xs, np.float32))
        dt = self.upload(np.ascontiguousarray(ts, np.float32))
        g = np.empty(CDF_ROOT_PARAMS, np.float32)
        try:
            self._check(self.lib.ocp_cdf_grad(self.h, dx.ptr, dt.ptr, int(dx.size), _ptr(g)))
        finally:
            dx.free()
            dt.free()
        return g

    def cdf_buckets(self, keys: DeviceArray, n: int | None = None, mode: int = 1) -> np.ndarray:
        n = keys.size if n is None else n
        q = self.empty(max(n, 1), np.uint32)
        try:
            self._check(self.lib.ocp_cdf_buckets(self.h, keys.ptr, int(n), int(mode), q.ptr))
            return q.to_host(n)
        finally:
            q.free()

    def cdf_boundaries(self) -> np.ndarray:
        fb = np.empty(CDF_BUCKETS, np.uint32)
        self._check(self.lib.ocp_cdf_boundaries(self.h, _ptr(fb)))
        return fb

    def sort_learned(self, keys: DeviceArray, keys_alt: DeviceArray, n: int | None = None) -> dict:
        n = keys.size if n is None else n
        st = LearnedStats()
        self._check(self.lib.ocp_sort_learned(self.h, keys.ptr, keys_alt.ptr, int(n),
                                              ctypes.byref(st)))
        return st.as_dict()


class _SlotContext(Context):
    """A pipeline slot's context: the Context API over a handle the pipeline
    owns (closing it frees only the arrays allocated through it)."""

    def __init__(self, lib, handle):
        self.lib = lib
        self.h = ctypes.c_void_p(handle)
        self._owned = []

    def close(self):
        for a in self._owned:
            a.free()
        self._owned.clear()
        self.h = None


class Pipeline:
    """A stream of frames built as by Context.build, consecutive frames
    overlapping on the device (ocp_ppl_*). Frames rotate over `depth` slots;
    each slot has its own stream, outputs, pinned host buffer and device
    points buffer of max_points float4."""

    def __init__(self, depth: int = 2, max_points: int = 0, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.ocp_ppl_create(ctypes.byref(h), device, int(depth), int(max_points))
        if rc != OCP_OK:
            raise OcpError(f"ocp_ppl_create failed with status {rc}")
        self.h = h
        self.depth = int(depth)
        self.max_points = int(max_points)
        self._slots = []
        for i in range(self.depth):
            c = ctypes.c_void_p()
            self._check(self.lib.ocp_ppl_context(self.h, i, ctypes.byref(c)))
            self._slots.append(_SlotContext(self.lib, c.value))

    def _check(self, rc: int):
        if rc != OCP_OK:
            msg = self.lib.ocp_ppl_last_error(self.h)
            raise OcpError(f"status {rc}: {msg.decode() if msg else ''}")

    def host_buffer(self, slot: int, n: int | None = None, float3: bool = False) -> np.ndarray:
        """The slot's pinned host buffer as an (n, 4) (or (n, 3)) float32 view,
        once its last upload has finished."""
        ptr = ctypes.c_void_p()
        self._check(self.lib.ocp_ppl_host_buffer(self.h, int(slot), ctypes.byref(ptr)))
        n = self.max_points if n is None else int(n)
        w = 3 if float3 else 4
        buf = (ctypes.c_float * (n * w)).from_address(ptr.value)
        return np.ctypeslib.as_array(buf).reshape(n, w)

    def submit(self, slot: int, points: DeviceArray | None, n: int, min_coord: float = 0.0,
               range_: float = 1024.0, float3: bool = False):
        """Enqueue one frame (points None: upload n points from the slot's
        host buffer first)."""
        if points is not None:
            float3 = points.shape[-1] == 3
        self._check(self.lib.ocp_ppl_submit(self.h, int(slot),
                                            points.ptr if points is not None else None,
                                            int(float3), int(n), float(min_coord),
                                            float(range_)))

    def wait(self, slot: int):
        self._check(self.lib.ocp_ppl_wait(self.h, int(slot)))

    def context(self, slot: int) -> Context:
        return self._slots[slot]

    def close(self):
        if getattr(self, "h", None):
            for c in self._slots:
                c.close()
            self.lib.ocp_ppl_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

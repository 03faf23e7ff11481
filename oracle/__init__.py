"""CPU oracle for the octree pipeline (ctypes over oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker or the timed CPU baseline. The product path (``pipeline_amd``) never
imports it.

Every function restates the reference CPU stage it names (kernels/0N_*.cpp,
include/shared/morton.h, include/shared/oct.h of ucsc-redwood/pipeline); the
C++ source is ``oracle/octree_oracle.cpp``. ``build_plain`` chains the stages
in the order of apps/cpu_demo/main.cpp:63-155 with the canonical call
convention: BRT over M = U-1 nodes, edge count and scan over M entries, node
slots K = max(1, max_i P[i]+e[i]).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libocref.so")

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

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F = ctypes.c_float
_INT = ctypes.c_int

_lib = None
_ref = None


def build() -> None:
    """Compile the oracle (and oracle/_ref where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `make -C oracle` first"
            )
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_split_by_3": (ctypes.c_uint32, [ctypes.c_uint32]),
            "orc_encode": (ctypes.c_uint32, [_F, _F, _F, _F, _F]),
            "orc_third_bits": (ctypes.c_uint32, [ctypes.c_uint32]),
            "orc_decode": (None, [ctypes.c_uint32, _F, _F, _P]),
            "orc_delta": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
            "orc_log2_ceil": (_I32, [ctypes.c_uint32]),
            "orc_morton": (None, [_P, _P, _I64, _F, _F, _INT]),
            "orc_morton3": (None, [_P, _P, _I64, _F, _F]),
            "orc_radix_sort": (None, [_P, _I64, _INT]),
            "orc_count_unique": (_I32, [_P, _I32]),
            "orc_build_radix_tree": (None, [_I32, _P, _P, _P, _P, _P, _P, _INT]),
            "orc_edge_count": (None, [_P, _P, _P, _I32, _INT]),
            "orc_partial_sum": (None, [_P, _I32, _P]),
            "orc_oct_alloc_count": (_I32, [_P, _P, _I32, _I32]),
            "orc_root_init": (None, [_P, _P, _P, _I32, _F, _F]),
            "orc_make_oct_nodes": (
                None, [_P, _P, _P, _P, _P, _P, _F, _F, _I32, _I32, _INT]),
            "orc_link_leaf_nodes": (
                None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _INT]),
            "orc_time_pipeline": (
                _I32, [_P, _I64, _F, _F, _INT, _INT, _P, _P]),
            "orc_max_threads": (_INT, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own CPU stages (oracle/_ref), or None where not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB_PATH):
            return None
        L = ctypes.CDLL(REF_LIB_PATH)
        sig = {
            "ref_simple_radix_sort": (None, [_P, _INT, _INT]),
            "ref_count_unique": (_INT, [_P, _INT]),
            "ref_build_radix_tree": (None, [_INT, _P, _P, _P, _P, _P, _P]),
            "ref_edge_count": (None, [_P, _P, _P, _INT]),
            "ref_partial_sum": (None, [_P, _INT, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ref = L
    return _ref


# ---------------------------------------------------------------- generator
def gpu_random_vec4(n: int, min_coord: float, range_: float, seed: int) -> np.ndarray:
    """kernels/cuda/00_init.cu:27-50 launched with at least n threads
    (apps/gpu_demo/main.cu:167-173): point i takes one hybrid LCG-Tausworthe
    step from z_k = (4i + k) * seed (mod 2^32); coordinates are
    min + ((x & 0xffffff) / 2^24) * range in float32, each product and sum
    rounded separately. Not a parity item of the pipeline (it only feeds it);
    parity unpinned beyond this restatement of the published algorithm."""
    i = np.arange(n, dtype=np.uint64)
    sd = np.uint64(np.uint32(np.int64(seed) & 0xffffffff))
    m32 = np.uint64(0xffffffff)
    z = [((i * np.uint64(4) + np.uint64(k)) * sd) & m32 for k in range(4)]

    def taus(v, s1, s2, s3, msk):
        return ((((v & np.uint64(msk)) << np.uint64(s3)) & m32) ^
                ((((v << np.uint64(s1)) & m32) ^ v) >> np.uint64(s2)))
    z1 = taus(z[0], 13, 19, 12, 4294967294)
    z2 = taus(z[1], 2, 25, 4, 4294967288)
    z3 = taus(z[2], 3, 11, 17, 4294967280)
    z4 = (z[3] * np.uint64(1664525) + np.uint64(1013904223)) & m32
    mn = np.float32(min_coord)
    rg = np.float32(range_)

    def unit(x):
        return (x & np.uint64(0xffffff)).astype(np.float32) / np.float32(16777216.0)
    out = np.empty((n, 4), np.float32)
    for c, x in enumerate((z1 ^ z2, z2 ^ z3, z3 ^ z4, z1 ^ z4)):
        out[:, c] = mn + unit(x) * rg
    return out


# ---------------------------------------------------------------- scalars
def encode(x, y, z, min_coord=0.0, range_=1024.0) -> int:
    return int(lib().orc_encode(x, y, z, min_coord, range_))


def decode(code, min_coord=0.0, range_=1024.0) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().orc_decode(code, min_coord, range_, _ptr(out))
    return out


def delta(a, b) -> int:
    return int(lib().orc_delta(a, b))


def log2_ceil(x) -> int:
    return int(lib().orc_log2_ceil(x))


# ---------------------------------------------------------------- stages
def morton(points: np.ndarray, min_coord=0.0, range_=1024.0, nthreads=1):
    pts = np.ascontiguousarray(points, np.float32)
    n = pts.shape[0]
    keys = np.zeros(n, np.uint32)
    if pts.shape[1] == 4:
        lib().orc_morton(_ptr(pts), _ptr(keys), n, min_coord, range_, nthreads)
    elif pts.shape[1] == 3:
        lib().orc_morton3(_ptr(pts), _ptr(keys), n, min_coord, range_)
    else:
        raise ValueError("points must be (n,4) or (n,3) float32")
    return keys


def radix_sort(keys: np.ndarray, nthreads=1) -> np.ndarray:
    k = np.array(keys, np.uint32, copy=True)
    lib().orc_radix_sort(_ptr(k), k.size, nthreads)
    return k


def count_unique(sorted_keys: np.ndarray):
    """Returns (unique keys, U); the reference works in place (03_unique.cpp)."""
    k = np.array(sorted_keys, np.uint32, copy=True)
    u = lib().orc_count_unique(_ptr(k), k.size)
    return k[:u].copy(), int(u)


def build_radix_tree(codes: np.ndarray, n_brt: int | None = None):
    """BRT over n_brt nodes (default U-1) of the sorted unique codes."""
    codes = np.ascontiguousarray(codes, np.uint32)
    n = codes.size - 1 if n_brt is None else n_brt
    n = max(n, 0)
    t = {
        "prefix_n": np.zeros(n, np.uint8),
        "has_leaf_left": np.zeros(n, np.uint8),
        "has_leaf_right": np.zeros(n, np.uint8),
        "left_child": np.zeros(n, np.int32),
        "parent": np.zeros(n, np.int32),
    }
    lib().orc_build_radix_tree(
        n, _ptr(codes), _ptr(t["prefix_n"]), _ptr(t["has_leaf_left"]),
        _ptr(t["has_leaf_right"]), _ptr(t["left_child"]), _ptr(t["parent"]), 1)
    return t


def edge_count(prefix_n, parent):
    n = prefix_n.size
    e = np.zeros(n, np.int32)
    lib().orc_edge_count(_ptr(np.ascontiguousarray(prefix_n, np.uint8)),
                         _ptr(np.ascontiguousarray(parent, np.int32)),
                         _ptr(e), n, 1)
    return e


def partial_sum(e):
    e = np.ascontiguousarray(e, np.int32)
    out = np.zeros(e.size, np.int32)
    lib().orc_partial_sum(_ptr(e), e.size, _ptr(out))
    return out


def octree(codes, tree, e, p, min_coord=0.0, range_=1024.0):
    """Root init + make + link; returns (nodes, K)."""
    codes = np.ascontiguousarray(codes, np.uint32)
    u = codes.size
    m = e.size
    k = int(lib().orc_oct_alloc_count(_ptr(p), _ptr(e), m, u))
    nodes = np.zeros(max(k, 0), OCT_NODE_DTYPE)
    if u == 0:
        return nodes, 0
    L = lib()
    L.orc_root_init(_ptr(nodes), _ptr(codes), _ptr(tree["prefix_n"]), m,
                    min_coord, range_)
    L.orc_make_oct_nodes(_ptr(nodes), _ptr(p), _ptr(e), _ptr(codes),
                         _ptr(tree["prefix_n"]), _ptr(tree["parent"]),
                         min_coord, range_, m, k, 1)
    L.orc_link_leaf_nodes(_ptr(nodes), _ptr(p), _ptr(e), _ptr(codes),
                          _ptr(tree["has_leaf_left"]),
                          _ptr(tree["has_leaf_right"]),
                          _ptr(tree["prefix_n"]), _ptr(tree["parent"]),
                          _ptr(tree["left_child"]), m, k, 1)
    return nodes, k


def build_plain(points: np.ndarray, min_coord=0.0, range_=1024.0):
    """All seven stages on one cloud; returns a dict of every stage output."""
    keys = morton(points, min_coord, range_)
    sorted_keys = np.sort(keys, kind="stable")
    ukeys, u = count_unique(sorted_keys)
    tree = build_radix_tree(ukeys)
    e = edge_count(tree["prefix_n"], tree["parent"])
    p = partial_sum(e)
    if p.size:
        p[0] = 0
    nodes, k = octree(ukeys, tree, e, p, min_coord, range_)
    return {
        "morton": keys, "sorted": sorted_keys, "unique": ukeys, "n_unique": u,
        **tree, "edge_count": e, "offsets": p, "nodes": nodes, "n_oct": k,
    }


def time_pipeline(points: np.ndarray, min_coord=0.0, range_=1024.0,
                  nthreads=0, reps=3):
    """Median per-stage ms of the whole chain on nthreads host threads
    (0 = all). Returns (dict of ms, U, K)."""
    pts = np.ascontiguousarray(points, np.float32)
    ms = np.zeros(8, np.float64)
    kk = np.zeros(1, np.int32)
    u = lib().orc_time_pipeline(_ptr(pts), pts.shape[0], min_coord, range_,
                                nthreads, reps, _ptr(ms), _ptr(kk))
    names = ["morton", "sort", "unique", "radix_tree", "edge_count",
             "prefix_sum", "octree", "total"]
    return dict(zip(names, ms.tolist())), int(u), int(kk[0])


def max_threads() -> int:
    return int(lib().orc_max_threads())

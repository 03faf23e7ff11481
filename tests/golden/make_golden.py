"""Writes tests/golden/golden.npz: inputs and expected stage outputs.

Run in the build container, where oracle/_ref/libocref.so (the reference's
own kernels/02_sort.cpp, 03_unique.cpp, 04_radix_tree.cpp, 05_edge_count.cpp
and 06_partial_sum.cpp, compiled by oracle/Makefile) exists:

    make -C oracle && python tests/golden/make_golden.py

Sort, unique, radix tree, edge count and prefix sum outputs come from the
reference's compiled functions (single thread, zeroed buffers, BRT over U-1
nodes). Morton codes and octree nodes come from the oracle, because the
reference sources for those stages need a header the image lacks (glm); the
tests pin them separately (closed forms, hand-traced cases).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from pipeline_amd import clouds  # noqa: E402


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ref_chain(codes_unsorted):
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libocref.so is not built (make -C oracle)")
    k = np.array(codes_unsorted, np.uint32)
    R.ref_simple_radix_sort(_p(k), k.size, 1)
    sorted_keys = k.copy()
    u = R.ref_count_unique(_p(k), k.size)
    ukeys = k[:u].copy()
    m = max(u - 1, 0)
    t = {n: np.zeros(m, d) for n, d in [("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
                                        ("has_leaf_right", np.uint8), ("left_child", np.int32),
                                        ("parent", np.int32)]}
    if m >= 2:
        R.ref_build_radix_tree(m, _p(ukeys), _p(t["prefix_n"]), _p(t["has_leaf_left"]),
                               _p(t["has_leaf_right"]), _p(t["left_child"]), _p(t["parent"]))
    else:  # U <= 2 reaches clz(0) in the reference: the canonical resolution
        t = oracle.build_radix_tree(ukeys)
    e = np.zeros(m, np.int32)
    p = np.zeros(m, np.int32)
    if m >= 1:
        R.ref_edge_count(_p(t["prefix_n"]), _p(t["parent"]), _p(e), m)
        R.ref_partial_sum(_p(e), m, _p(p))
        p[0] = 0
    return sorted_keys, ukeys, u, t, e, p


def case_from_points(name, pts, out):
    codes = oracle.morton(pts)
    sorted_keys, ukeys, u, t, e, p = ref_chain(codes)
    nodes, k = oracle.octree(ukeys, t, e, p)
    out[f"{name}/points"] = pts
    out[f"{name}/morton"] = codes
    out[f"{name}/sorted"] = sorted_keys
    out[f"{name}/unique"] = ukeys
    for kk, vv in t.items():
        out[f"{name}/{kk}"] = vv
    out[f"{name}/edge_count"] = e
    out[f"{name}/offsets"] = p
    out[f"{name}/nodes"] = nodes.view(np.uint8).reshape(-1, 60) if k else np.zeros((0, 60), np.uint8)


def point_of_code(code):
    """A point at the centre of the unit cell of a 30-bit code in [0,1024)^3."""
    x = oracle.lib().orc_third_bits(code)
    y = oracle.lib().orc_third_bits(code >> 1)
    z = oracle.lib().orc_third_bits(code >> 2)
    return [x + 0.5, y + 0.5, z + 0.5, 0.0]


def main():
    out = {}
    case_from_points("uniform64", clouds.uniform(64, seed=1), out)
    case_from_points("uniform4096", clouds.uniform(4096, seed=2), out)
    case_from_points("uniform8000", clouds.uniform(8000, seed=3), out)
    case_from_points("clustered8000", clouds.clustered(8000, seed=4), out)
    case_from_points("one_point", np.array([[3.0, 4.0, 5.0, 0.0]], np.float32), out)
    case_from_points("two_points", np.array([[1.0, 1.0, 1.0, 0.0], [900.0, 2.0, 3.0, 0.0]],
                                            np.float32), out)
    case_from_points("three_points", np.array([[1.0, 1.0, 1.0, 0.0], [2.5, 1.0, 1.0, 0.0],
                                               [700.0, 800.0, 900.0, 0.0]], np.float32), out)
    # a lone outlier as the last key (the short-range root of
    # 04_radix_tree.cpp:63-65 then differs from the true root)
    lo = clouds.uniform(3000, seed=5)
    lo[:, :3] *= np.float32(0.4)
    lo[-1, :3] = [1023.5, 1023.5, 1023.5]
    case_from_points("lone_outlier", lo, out)
    # sibling pairs that differ only in bit 0 (leaf level 11, shift -3)
    pairs = []
    for base in range(0, 1 << 30, (1 << 30) // 97):
        b = base & ~1
        pairs += [point_of_code(b), point_of_code(b | 1)]
    case_from_points("depth10_pairs", np.array(pairs, np.float32), out)
    # chains longer than one step next to single steps
    chain = [point_of_code(c) for c in (0, 1, 1 << 15, (1 << 15) + 7, 1 << 27, (1 << 27) + 1,
                                        (1 << 29) + 5, (1 << 29) + (1 << 20))]
    case_from_points("long_chains", np.array(chain, np.float32), out)
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()

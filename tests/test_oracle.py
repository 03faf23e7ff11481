"""Pins the CPU oracle (oracle/) - CPU only.

Each stage is checked against something that fixes it independently of the
oracle's own code:
  * the reference's compiled CPU stages (oracle/_ref: kernels/02_sort.cpp,
    03_unique.cpp, 04_radix_tree.cpp, 05_edge_count.cpp, 06_partial_sum.cpp)
    where they are built, and the golden fixtures written from them;
  * closed forms and brute force (Morton bit interleave, float32 arithmetic
    in numpy, np.sort/np.unique/np.cumsum, a linear-scan Karras tree);
  * hand-traced known answers of the octree stage (include/shared/oct.h),
    whose reference source needs a header this image lacks.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from octree_checks import check_octree
from pipeline_amd import clouds

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "golden.npz")
F32 = np.float32


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ Morton
def test_morton_known_values():
    assert oracle.encode(1.0, 0.0, 0.0) == 1
    assert oracle.encode(0.0, 1.0, 0.0) == 2
    assert oracle.encode(0.0, 0.0, 1.0) == 4
    assert oracle.encode(1023.0, 1023.0, 1023.0) == 0x3FFFFFFF
    assert oracle.encode(1023.9999, 0.0, 0.0) == oracle.encode(1023.0, 0.0, 0.0)
    # (a - min) / range * 1024 == 1024 wraps to 0 (& 0x3ff, morton.h:23)
    assert oracle.encode(1024.0, 1024.0, 1024.0) == 0


def test_split_by_3_is_a_bit_interleave():
    L = oracle.lib()
    for a in range(1024):
        brute = 0
        for b in range(10):
            brute |= ((a >> b) & 1) << (3 * b)
        assert L.orc_split_by_3(a) == brute
        assert L.orc_third_bits(brute) == a
    rng = np.random.default_rng(0)
    for m in rng.integers(0, 1 << 32, 2000, dtype=np.uint64):
        m = int(m)
        brute = 0
        for b in range(10):
            brute |= ((m >> (3 * b)) & 1) << b
        assert L.orc_third_bits(m) == brute


def _np_quantize(x, mn, rg):
    q = ((x.astype(F32) - F32(mn)) / F32(rg)) * F32(1024)
    return (np.trunc(q).astype(np.int64) & 0x3FF).astype(np.uint32)


def _np_interleave(q):
    out = np.zeros(q.shape[0], np.uint64)
    for b in range(10):
        for axis in range(3):
            out |= ((q[:, axis].astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + axis)
    return out.astype(np.uint32)


@pytest.mark.parametrize("mn,rg", [(0.0, 1024.0), (-3.5, 77.25), (100.0, 0.5)])
def test_morton_matches_float32_arithmetic(mn, rg):
    rng = np.random.default_rng(7)
    pts = (rng.random((20000, 3)) * rg + mn).astype(F32)
    # values next to quantisation boundaries
    k = rng.integers(0, 1024, (4000, 3))
    edge = (F32(mn) + (k.astype(F32) / F32(1024)) * F32(rg)).astype(F32)
    pts = np.concatenate([pts, edge, np.nextafter(edge, F32(-1e30)), np.nextafter(edge, F32(1e30))])
    pts = np.clip(pts, F32(mn), np.nextafter(F32(mn + rg), F32(-1e30)))
    q = np.stack([_np_quantize(pts[:, a], mn, rg) for a in range(3)], 1)
    np.testing.assert_array_equal(oracle.morton(pts, mn, rg), _np_interleave(q))
    p4 = np.zeros((pts.shape[0], 4), F32)
    p4[:, :3] = pts
    p4[:, 3] = 123.0  # w is ignored
    np.testing.assert_array_equal(oracle.morton(p4, mn, rg), _np_interleave(q))


@pytest.mark.parametrize("mn,rg", [(0.0, 1024.0), (-3.5, 77.25)])
def test_decode_matches_float32_arithmetic(mn, rg):
    rng = np.random.default_rng(8)
    for code in rng.integers(0, 1 << 30, 500):
        code = int(code)
        got = oracle.decode(code, mn, rg)
        L = oracle.lib()
        q = np.array([L.orc_third_bits(code >> a) for a in range(3)], np.uint32)
        exp = (q.astype(F32) / F32(1024.0)) * F32(rg) + F32(mn)  # separately rounded
        np.testing.assert_array_equal(got[:3], exp.astype(F32))
        assert got[3] == F32(1.0)


# ------------------------------------------------------------------ sort
def test_sort_reference_kat():
    # tests/test_simple_radix_sort_cpu.cpp with tests/xmake.lua:9-13 runargs
    keys = np.arange(10, 0, -1, dtype=np.uint32)
    np.testing.assert_array_equal(oracle.radix_sort(keys, nthreads=4), np.arange(1, 11))


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_sort_matches_numpy(nthreads):
    rng = np.random.default_rng(nthreads)
    for n in (1, 2, 255, 256, 257, 100003):
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        np.testing.assert_array_equal(oracle.radix_sort(keys, nthreads), np.sort(keys))


# ------------------------------------------------------------------ unique / scan / edge
def test_unique_and_partial_sum_closed_forms():
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 500, 5000, dtype=np.uint32))
    u, n = oracle.count_unique(keys)
    np.testing.assert_array_equal(u, np.unique(keys))
    assert n == np.unique(keys).size
    assert oracle.count_unique(np.zeros(0, np.uint32))[1] == 0
    v = rng.integers(-100, 100, 7777, dtype=np.int32)
    np.testing.assert_array_equal(oracle.partial_sum(v), np.cumsum(v).astype(np.int32))


def test_edge_count_closed_form():
    codes = np.unique(np.random.default_rng(4).integers(0, 1 << 30, 30000, dtype=np.uint32))
    t = oracle.build_radix_tree(codes)
    e = oracle.edge_count(t["prefix_n"], t["parent"])
    pn = t["prefix_n"].astype(np.int32)
    exp = pn // 3 - pn[t["parent"]] // 3
    exp[0] = 0
    np.testing.assert_array_equal(e, exp)


# ------------------------------------------------------------------ radix tree
def _delta(a, b):
    x = int(a) ^ int(b)
    return ((32 - x.bit_length()) - 1) & 0xFF


def _linear_karras(codes):
    """04_radix_tree.cpp's node function with linear scans in place of the
    exponential and binary searches, root range [0, n-1], parents by
    sequential last writer on zeroed memory."""
    n = codes.size - 1
    pn = np.zeros(n, np.uint8)
    hl = np.zeros(n, np.uint8)
    hr = np.zeros(n, np.uint8)
    lc = np.zeros(n, np.int32)
    par = np.zeros(n, np.int32)
    for i in range(n):
        if i == 0:
            d, j = 1, n - 1
        else:
            diff = _delta(codes[i], codes[i + 1]) - _delta(codes[i], codes[i - 1])
            d = (diff > 0) - (diff < 0)
            dmin = _delta(codes[i], codes[i - d])
            j = i
            while 0 <= j + d <= n and _delta(codes[i], codes[j + d]) > dmin:
                j += d
        l = abs(j - i)
        dn = _delta(codes[i], codes[j])
        s = 0
        for cand in range(l):  # largest s < l with delta(i, i + s*d) > dn
            if _delta(codes[i], codes[i + cand * d]) > dn:
                s = cand
        gamma = i + s * d + min(d, 0)
        pn[i], lc[i] = dn, gamma
        hl[i] = min(i, j) == gamma
        hr[i] = max(i, j) == gamma + 1
        if not hl[i] and 0 <= gamma < n:
            par[gamma] = i
        if not hr[i] and 0 <= gamma + 1 < n:
            par[gamma + 1] = i
    return {"prefix_n": pn, "has_leaf_left": hl, "has_leaf_right": hr,
            "left_child": lc, "parent": par}


@pytest.mark.parametrize("seed", range(6))
def test_radix_tree_matches_linear_scan_construction(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 300))
    if seed % 2:
        codes = np.unique(rng.integers(0, 1 << 30, n, dtype=np.uint32))
    else:  # clustered: many shared prefixes
        codes = np.unique((rng.integers(0, 64, n) << 20 | rng.integers(0, 1 << 8, n)).astype(np.uint32))
    if seed == 4:
        codes = np.concatenate([codes, [np.uint32((1 << 30) - 1)]]).astype(np.uint32)
    exp = _linear_karras(codes)
    got = oracle.build_radix_tree(codes)
    for k in exp:
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)


def _ref_tree(codes):
    R = oracle.ref_lib()
    m = codes.size - 1
    t = {n: np.zeros(m, d) for n, d in [("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
                                        ("has_leaf_right", np.uint8), ("left_child", np.int32),
                                        ("parent", np.int32)]}
    R.ref_build_radix_tree(m, _p(codes), _p(t["prefix_n"]), _p(t["has_leaf_left"]),
                           _p(t["has_leaf_right"]), _p(t["left_child"]), _p(t["parent"]))
    return t


def test_stages_match_compiled_reference():
    # Looked up inside the test, not at collection: a GPU-only run never
    # maps the compiled reference.
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(11)
    for kind in range(3):
        n = 50000
        if kind == 0:
            keys = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        elif kind == 1:
            keys = (rng.integers(0, 16, n) << 25 | rng.integers(0, 1 << 12, n)).astype(np.uint32)
        else:
            keys = np.concatenate([rng.integers(0, 1 << 20, n - 1), [(1 << 30) - 1]]).astype(np.uint32)
        k = keys.copy()
        R.ref_simple_radix_sort(_p(k), k.size, 4)
        np.testing.assert_array_equal(oracle.radix_sort(keys, 4), k)
        kk = k.copy()
        u = R.ref_count_unique(_p(kk), kk.size)
        ou, on = oracle.count_unique(k)
        assert on == u
        codes = kk[:u].copy()
        np.testing.assert_array_equal(ou, codes)
        ref_t = _ref_tree(codes)
        t = oracle.build_radix_tree(codes)
        for name in ref_t:
            np.testing.assert_array_equal(t[name], ref_t[name], err_msg=name)
        m = u - 1
        e = np.zeros(m, np.int32)
        R.ref_edge_count(_p(ref_t["prefix_n"]), _p(ref_t["parent"]), _p(e), m)
        np.testing.assert_array_equal(oracle.edge_count(t["prefix_n"], t["parent"]), e)
        p = np.zeros(m, np.int32)
        R.ref_partial_sum(_p(e), m, _p(p))
        np.testing.assert_array_equal(oracle.partial_sum(e), p)


# ------------------------------------------------------------------ octree
def _node(children, corner, cell, nmask, lmask=0):
    n = np.zeros(1, oracle.OCT_NODE_DTYPE)
    n["children"][0] = children
    n["corner"][0] = corner
    n["cell_size"][0] = cell
    n["child_node_mask"][0] = nmask
    n["child_leaf_mask"][0] = lmask
    return n


def test_octree_hand_traced_chain():
    """codes 0, 1, 64, 2^29, traced by hand through 04_radix_tree.cpp,
    05_edge_count.cpp, 06_partial_sum.cpp, apps/cpu_demo/main.cpp:125-132 and
    oct.h:24-134 (the trace is in the comments)."""
    codes = np.array([0, 1, 64, 1 << 29], np.uint32)
    t = oracle.build_radix_tree(codes)
    # node 0: short-range root [0, 2], delta(0,64)=24, split after key 1
    # node 1: d=-1, range [0,1], delta 30, both children leaves
    # node 2: d=-1, range [0,2], delta 24, split at 1; writes parent[1]=2
    np.testing.assert_array_equal(t["prefix_n"], [24, 30, 24])
    np.testing.assert_array_equal(t["has_leaf_left"], [0, 1, 0])
    np.testing.assert_array_equal(t["has_leaf_right"], [1, 1, 1])
    np.testing.assert_array_equal(t["left_child"], [1, 0, 1])
    np.testing.assert_array_equal(t["parent"], [0, 2, 0])
    e = oracle.edge_count(t["prefix_n"], t["parent"])
    np.testing.assert_array_equal(e, [0, 2, 0])       # 30/3 - 24/3 = 2
    p = oracle.partial_sum(e)
    p[0] = 0
    np.testing.assert_array_equal(p, [0, 2, 2])       # inclusive scan
    nodes, k = oracle.octree(codes, t, e, p)
    assert k == 4                                     # max(P + e) = 2 + 2
    # chain of node 1 at slots 2 (level 10) and 3 (top, level 9); its top link
    # walks 2 -> 0 -> 0 ... and lands in slot P[0] = 0 as child 0; the link
    # pass then overwrites root child 0 with leaf 2 (node 0's right leaf,
    # level 9, child (64 >> 3) & 7 = 0) and node 1's leaves 0 and 1 reach
    # slot 2 child (code >> ((30 - 33) & 31)) & 7 = 0, the last being 1.
    exp = np.concatenate([
        _node([2, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1], 1024.0, 1),
        _node([0] * 8, [0, 0, 0, 0], 0.0, 0),              # slot 1: never written
        _node([1, 0, 0, 0, 0, 0, 0, 0], [1, 0, 0, 1], 256.0, 0),
        _node([0, 2, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1], 512.0, 2),
    ])
    np.testing.assert_array_equal(nodes.view(np.uint8), exp.view(np.uint8))


def test_octree_hand_traced_short_root():
    """codes 0, 8, 2^29: the short-range root never sees the last key, both
    nodes have e = 0, every leaf link lands in the root slot."""
    codes = np.array([0, 8, 1 << 29], np.uint32)
    t = oracle.build_radix_tree(codes)
    np.testing.assert_array_equal(t["prefix_n"], [27, 27])
    np.testing.assert_array_equal(t["left_child"], [0, 0])
    np.testing.assert_array_equal(t["parent"], [0, 0])
    e = oracle.edge_count(t["prefix_n"], t["parent"])
    p = oracle.partial_sum(e)
    p[0] = 0
    nodes, k = oracle.octree(codes, t, e, p)
    assert k == 1
    exp = _node([1, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1], 1024.0, 0)
    np.testing.assert_array_equal(nodes.view(np.uint8), exp.view(np.uint8))


def test_octree_hand_traced_overlapping_chains():
    """codes 0, 1, 512, 4096, 2^21: a chain of three slots overlapped by the
    next, shorter chain, traced by hand through 04_radix_tree.cpp:39-121,
    05_edge_count.cpp:3-16, 06_partial_sum.cpp:3-15, the root set-up of
    apps/cpu_demo/main.cpp:125-132 and oct.h:24-134. Pins the per-step level
    of the make loop (j = 0, 1), the last writer of a corner and cell size
    where chains overlap, a make-phase child that the link pass overwrites,
    the 31-step walk to the root and the leaf shift of level 11."""
    codes = np.array([0, 1, 512, 4096, 1 << 21], np.uint32)
    t = oracle.build_radix_tree(codes)
    # delta: (c0,c1)=30, (c0|c1,c2)=21, (c0..c2,c3)=18, (*,c4)=9.
    # node 0: root quirk, range [0, 3], delta 18, split after key 2 (delta(c0,c2)
    #   = 21 > 18, delta(c0,c3) = 18 is not): gamma 2, right leaf 3, parent[2]=0
    # node 1: d = -1, range [0, 1], delta 30, both children leaves 0 and 1
    # node 2: d = -1, range [0, 2], delta 21, gamma 1 (delta(c2,c1) = 21 is not
    #   > 21): left child node 1 (parent[1] = 2), right leaf 2
    # node 3: d = -1, range [0, 3], delta 18, gamma 2: left child node 2
    #   (parent[2] = 3 overwrites node 0's write), right leaf 3
    np.testing.assert_array_equal(t["prefix_n"], [18, 30, 21, 18])
    np.testing.assert_array_equal(t["has_leaf_left"], [0, 1, 0, 0])
    np.testing.assert_array_equal(t["has_leaf_right"], [1, 1, 1, 1])
    np.testing.assert_array_equal(t["left_child"], [2, 0, 1, 2])
    np.testing.assert_array_equal(t["parent"], [0, 2, 3, 0])
    e = oracle.edge_count(t["prefix_n"], t["parent"])
    np.testing.assert_array_equal(e, [0, 3, 1, 0])   # 10-7, 7-6, 6-6
    p = oracle.partial_sum(e)
    p[0] = 0
    np.testing.assert_array_equal(p, [0, 3, 4, 4])   # chain 1: slots 3..5, chain 2: slot 4
    nodes, k = oracle.octree(codes, t, e, p)
    assert k == 6                                    # max(P + e) = 3 + 3
    # root (level 18/3 = 6): corner decode(0), cell 1024.
    # i = 1, j = 0: level 10, child (1 >> 0) & 7 = 1 -> slot 4 children[1] = 3;
    #   slot 3 corner decode(1) = (1,0,0), cell 1024 / 2^(10-6) = 64.
    # i = 1, j = 1: level 9, child (1 >> 3) & 7 = 0 -> slot 5 children[0] = 4;
    #   slot 4 corner (0,0,0), cell 128.
    # i = 1, top: parent[1] = 2 has e = 1, so slot P[2] = 4, top level
    #   10 - 3 + 1 = 8, child (1 >> 6) & 7 = 0 -> slot 4 children[0] = 5;
    #   slot 5 corner (0,0,0), cell 256. Slot 4 mask = 0b11.
    # i = 2, top: parent 3 (e = 0) -> 0 (e = 0) -> 0 ... capped at 31 steps:
    #   slot P[0] = 0, level 7, child (512 >> 9) & 7 = 1 -> slot 0 children[1]
    #   = 4; slot 4 corner decode(512) = (8,0,0), cell 512: chain 2 is the last
    #   writer of slot 4's geometry.
    # links: node 0 right leaf 3 at level 7: child (4096 >> 9) & 7 = 0, walk
    #   stops at the root -> slot 0 children[0] = 3. Node 1 leaves 0 and 1 at
    #   level 11: shift (30 - 33) & 31 = 29, child 0 -> slot 3 children[0] = 0,
    #   then = 1. Node 2 right leaf 2 at level 8: child (512 >> 6) & 7 = 0 ->
    #   slot 4 children[0] = 2, overwriting the make-phase 5. Node 3 right leaf
    #   3 -> slot 0 children[0] = 3 again. Slots 1 and 2: never written.
    exp = np.concatenate([
        _node([3, 4, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1], 1024.0, 0b10),
        _node([0] * 8, [0, 0, 0, 0], 0.0, 0),
        _node([0] * 8, [0, 0, 0, 0], 0.0, 0),
        _node([1, 0, 0, 0, 0, 0, 0, 0], [1, 0, 0, 1], 64.0, 0),
        _node([2, 3, 0, 0, 0, 0, 0, 0], [8, 0, 0, 1], 512.0, 0b11),
        _node([4, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1], 256.0, 0b1),
    ])
    np.testing.assert_array_equal(nodes.view(np.uint8), exp.view(np.uint8))
    assert check_octree(nodes, codes, t["prefix_n"], e, p) == 4


@pytest.mark.parametrize("kind,n", [("uniform", 200000), ("clustered", 200000),
                                    ("uniform", 5), ("clustered", 4096)])
def test_oracle_octree_structure(kind, n):
    """The oracle's octree output satisfies the order-independent structure
    of the reference's semantics (tests/octree_checks.py)."""
    gen = clouds.uniform if kind == "uniform" else clouds.clustered
    r = oracle.build_plain(gen(n, seed=n + 3))
    written = check_octree(r["nodes"], r["unique"], r["prefix_n"], r["edge_count"],
                           r["offsets"])
    assert written >= 1


def test_octree_structure_check_catches_planted_mistakes():
    r = oracle.build_plain(clouds.clustered(20000, seed=9))
    args = (r["unique"], r["prefix_n"], r["edge_count"], r["offsets"])
    check_octree(r["nodes"], *args)
    written = np.nonzero(r["nodes"]["cell_size"] != 0)[0]
    s = int(written[len(written) // 2])
    bad = r["nodes"].copy()
    bad["corner"][s, 0] += np.float32(0.5)           # off the grid
    with pytest.raises(AssertionError):
        check_octree(bad, *args)
    bad = r["nodes"].copy()
    bad["cell_size"][s] *= np.float32(0.75)          # not range / 2^k
    with pytest.raises(AssertionError):
        check_octree(bad, *args)
    bad = r["nodes"].copy()
    bad[s] = np.zeros(1, bad.dtype)[0]               # a covered slot left zero
    with pytest.raises(AssertionError):
        check_octree(bad, *args)
    gaps = np.nonzero(r["nodes"]["cell_size"] == 0)[0]
    if gaps.size:
        bad = r["nodes"].copy()
        bad["children"][int(gaps[0]), 3] = 7         # a gap written
        with pytest.raises(AssertionError):
            check_octree(bad, *args)
    bad = r["nodes"].copy()
    bad["child_leaf_mask"][s] = 1
    with pytest.raises(AssertionError):
        check_octree(bad, *args)


def test_single_and_empty_clouds():
    r = oracle.build_plain(np.array([[3.0, 4.0, 5.0, 0.0]], np.float32))
    assert r["n_unique"] == 1 and r["n_oct"] == 1
    exp = _node([0] * 8, [0, 0, 0, 1], 1024.0, 0)      # root level 0 -> prefix 0
    np.testing.assert_array_equal(r["nodes"].view(np.uint8), exp.view(np.uint8))
    r = oracle.build_plain(np.zeros((0, 4), np.float32))
    assert r["n_unique"] == 0 and r["n_oct"] == 0


# ------------------------------------------------------------------ golden
@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _cases(g):
    return sorted({k.split("/")[0] for k in g.files})


def test_golden_fixtures_reproduced(golden):
    for name in _cases(golden):
        g = lambda k: golden[f"{name}/{k}"]  # noqa: E731
        r = oracle.build_plain(g("points"))
        np.testing.assert_array_equal(r["morton"], g("morton"), err_msg=name)
        np.testing.assert_array_equal(r["sorted"], g("sorted"), err_msg=name)
        np.testing.assert_array_equal(r["unique"], g("unique"), err_msg=name)
        for k in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
                  "edge_count", "offsets"):
            np.testing.assert_array_equal(r[k], g(k), err_msg=f"{name}/{k}")
        np.testing.assert_array_equal(r["nodes"].view(np.uint8).reshape(-1, 60), g("nodes"),
                                      err_msg=f"{name}/nodes")


def test_multithreaded_variants_match_single_thread():
    rng = np.random.default_rng(12)
    pts = np.zeros((60000, 4), np.float32)
    pts[:, :3] = rng.random((60000, 3), dtype=np.float32) * 1024
    k1 = oracle.morton(pts, nthreads=1)
    np.testing.assert_array_equal(oracle.morton(pts, nthreads=8), k1)
    codes = np.unique(k1)
    t1 = oracle.build_radix_tree(codes)
    m = codes.size - 1
    t8 = {n: np.zeros(m, d) for n, d in [("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
                                         ("has_leaf_right", np.uint8), ("left_child", np.int32),
                                         ("parent", np.int32)]}
    oracle.lib().orc_build_radix_tree(m, _p(codes), *[_p(t8[n]) for n in t8], 8)
    for n in t1:
        np.testing.assert_array_equal(t8[n], t1[n], err_msg=n)

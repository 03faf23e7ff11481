"""GPU parity of every per-stage entry point against the CPU oracle.

Each test calls the HIP path through the C ABI (pipeline_amd.Context) and
compares bit for bit with the oracle (oracle/, the reference CPU path
restated) on the same seeded inputs. Sizes cover tile boundaries of every
kernel (OneSweep 7680, unique/scan 4096, node tiles 256) and the edge cases
the reference code reaches (U = 1, 2, 3, duplicates, a lone outlier key).
"""
import numpy as np
import pytest

import oracle
from pipeline_amd import OCT_NODE_DTYPE, clouds

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097, 7679, 7680,
         7681, 65536, 100003]


def _pts(n, seed, kind="uniform"):
    gen = clouds.uniform if kind == "uniform" else clouds.clustered
    return gen(n, seed=seed)


@pytest.mark.parametrize("n", [1, 5, 4095, 4096, 4097, 65537, 307200])
def test_morton_float4(ctx, n):
    p = _pts(n, 11 + n)
    d = ctx.upload(p)
    k = ctx.empty(n, np.uint32)
    ctx.k_ComputeMortonCode(d, k, n, 0.0, 1024.0)
    np.testing.assert_array_equal(k.to_host(), oracle.morton(p))


def test_morton_float3_and_boundaries(ctx):
    edge = np.array([0.0, 1e-30, 0.5, 1.0, 511.99997, 512.0, 1023.0,
                     1023.99994, 1024.0, 3.0e-3], np.float32)
    g = np.stack(np.meshgrid(edge, edge, edge, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(3)
    p3 = np.concatenate([g, rng.random((50000, 3), dtype=np.float32) * 1024]).astype(np.float32)
    d = ctx.upload(p3)
    k = ctx.empty(p3.shape[0], np.uint32)
    ctx.k_ComputeMortonCode(d, k, p3.shape[0], 0.0, 1024.0)
    np.testing.assert_array_equal(k.to_host(), oracle.morton(p3))
    # a box that is not [0, 1024): min -3.5, range 77.25
    p4 = np.zeros((p3.shape[0], 4), np.float32)
    p4[:, :3] = p3 * np.float32(77.25 / 1024.0) - np.float32(3.5)
    d4 = ctx.upload(p4)
    ctx.k_ComputeMortonCode(d4, k, p4.shape[0], -3.5, 77.25)
    np.testing.assert_array_equal(k.to_host(), oracle.morton(p4, -3.5, 77.25))


@pytest.mark.parametrize("n", SIZES + [1_000_003])
def test_sort_random(ctx, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 30, n, dtype=np.uint32)
    a = ctx.upload(keys)
    b = ctx.empty(n, np.uint32)
    ctx.k_SortKeys(a, b, n)
    np.testing.assert_array_equal(a.to_host(), np.sort(keys))


def test_sort_constant_digits_copy_through(ctx):
    """Passes whose digit is the same for every key copy the keys (and their
    values) through: the result is still the sort and the stable argsort."""
    rng = np.random.default_rng(9)
    cases = (np.uint32(0xAB120000) | rng.integers(0, 1 << 16, 300001).astype(np.uint32),
             (rng.integers(0, 256, 100000).astype(np.uint32) << np.uint32(8)) |
             np.uint32(0x5A0000AB))
    for keys in cases:
        n = keys.size
        a = ctx.upload(keys)
        b = ctx.empty(n, np.uint32)
        ctx.k_SortKeys(a, b, n)
        np.testing.assert_array_equal(a.to_host(), np.sort(keys))
        k = ctx.upload(keys)
        v = ctx.upload(np.arange(n, dtype=np.uint32))
        ka = ctx.empty(n, np.uint32)
        va = ctx.empty(n, np.uint32)
        ctx.k_SortPairs(k, v, ka, va, n)
        np.testing.assert_array_equal(k.to_host(), np.sort(keys))
        np.testing.assert_array_equal(v.to_host(), np.argsort(keys, kind="stable"))


def test_sort_kat_and_duplicates(ctx):
    # tests/test_simple_radix_sort_cpu.cpp + tests/xmake.lua:9-13: 10..1 -> 1..10
    kat = np.arange(10, 0, -1, dtype=np.uint32)
    a = ctx.upload(kat)
    b = ctx.empty(10, np.uint32)
    ctx.k_SortKeys(a, b, 10)
    np.testing.assert_array_equal(a.to_host(), np.arange(1, 11, dtype=np.uint32))
    # heavy duplicates, full 32-bit keys, constant input
    rng = np.random.default_rng(5)
    for keys in (rng.integers(0, 7, 200000, dtype=np.uint32),
                 rng.integers(0, 2**32, 123457, dtype=np.uint64).astype(np.uint32),
                 np.full(50000, 0xDEADBEEF, np.uint32)):
        a = ctx.upload(keys)
        b = ctx.empty(keys.size, np.uint32)
        ctx.k_SortKeys(a, b, keys.size)
        np.testing.assert_array_equal(a.to_host(), np.sort(keys))


def test_digit_binning_single_pass(ctx):
    rng = np.random.default_rng(9)
    n = 54321
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    a = ctx.upload(keys)
    h = ctx.empty(1024, np.uint32)
    ctx.k_GlobalHistogram(a, h, n)
    hist = h.to_host().reshape(4, 256)
    for p in range(4):
        np.testing.assert_array_equal(hist[p], np.bincount((keys >> (8 * p)) & 255, minlength=256))
    b = ctx.empty(n, np.uint32)
    for shift in (0, 8, 16, 24):
        ctx.k_DigitBinning(h, a, b, n, shift)
        d = (keys >> shift) & 255
        np.testing.assert_array_equal(b.to_host(), keys[np.argsort(d, kind="stable")])


@pytest.mark.parametrize("n", SIZES)
def test_unique(ctx, n):
    rng = np.random.default_rng(100 + n)
    keys = np.sort(rng.integers(0, max(n // 3, 1), n, dtype=np.uint32))
    a = ctx.upload(keys)
    u = ctx.k_CountUnique(a, n)
    exp, eu = oracle.count_unique(keys)
    assert u == eu
    np.testing.assert_array_equal(a.to_host()[:u], exp)


def _tree_on_gpu(ctx, codes):
    m = codes.size - 1
    c = ctx.upload(codes)
    out = {k: ctx.zeros(max(m, 1), dt) for k, dt in
           [("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
            ("has_leaf_right", np.uint8), ("left_child", np.int32), ("parent", np.int32)]}
    # poison parent: the kernel must write every entry it owns
    out["parent"].copy_from(np.full(max(m, 1), -7, np.int32))
    ctx.k_BuildRadixTree(m, c, out["prefix_n"], out["has_leaf_left"],
                         out["has_leaf_right"], out["left_child"], out["parent"])
    return c, {k: v.to_host()[:m] for k, v in out.items()}, out


def _unique_codes(n, seed, kind="uniform"):
    return np.unique(oracle.morton(_pts(n, seed, kind)))


TREE_CASES = {
    "u2": np.array([5, 9], np.uint32),
    "u3": np.array([1, 2, 1 << 29], np.uint32),
    "outlier": np.concatenate([np.arange(1000, dtype=np.uint32) * 3, [1 << 29]]).astype(np.uint32),
    "pairs": np.unique(np.array([0, 1, 8, 9, 64, 65, 512, 513, 4096, 4097], np.uint32)),
    "uniform": None,
    "clustered": None,
}


@pytest.mark.parametrize("case", list(TREE_CASES))
def test_radix_tree(ctx, case):
    codes = TREE_CASES[case]
    if codes is None:
        codes = _unique_codes(200000, 7, case)
    _, got, _ = _tree_on_gpu(ctx, codes)
    exp = oracle.build_radix_tree(codes)
    for k in exp:
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)


@pytest.mark.parametrize("n_codes", [2304, 2305, 2306, 2560, 2561, 2817, 3073])
def test_radix_tree_interior_tile_threshold(ctx, n_codes):
    """Tiles of 256 nodes at least 1024 nodes (twice the local search bound,
    half the 1024-code halo) from both ends of [0, n] skip the search bound
    checks. n = n_codes - 1 = 2304 is the first size with such a tile (nodes
    1024..1279); the sizes around it and around the next tiles put the last
    interior tile right at the threshold."""
    rng = np.random.default_rng(n_codes)
    codes = np.unique(rng.integers(0, 1 << 30, 3 * n_codes, dtype=np.int64).astype(np.uint32))
    codes = np.sort(rng.choice(codes, n_codes, replace=False))
    assert codes.size == n_codes
    _, got, _ = _tree_on_gpu(ctx, codes)
    exp = oracle.build_radix_tree(codes)
    for k in exp:
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_edge_count_and_scan(ctx, kind):
    codes = _unique_codes(150000, 21, kind)
    _, got, dev = _tree_on_gpu(ctx, codes)
    m = codes.size - 1
    e = ctx.empty(m, np.int32)
    ctx.k_EdgeCount(dev["prefix_n"], dev["parent"], e, m)
    exp_e = oracle.edge_count(got["prefix_n"], got["parent"])
    np.testing.assert_array_equal(e.to_host(), exp_e)
    p = ctx.empty(m, np.int32)
    ctx.k_PartialSum(e, 0, m, p)
    np.testing.assert_array_equal(p.to_host(), oracle.partial_sum(exp_e))


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 300007])
def test_partial_sum_generic(ctx, n):
    rng = np.random.default_rng(n)
    v = rng.integers(-50, 50, n, dtype=np.int32)
    d = ctx.upload(v)
    o = ctx.empty(n, np.int32)
    ctx.k_PartialSum(d, 0, n, o)
    np.testing.assert_array_equal(o.to_host(), np.cumsum(v, dtype=np.int64).astype(np.int32))


def _nodes_equal(got, exp):
    assert got.shape == exp.shape
    np.testing.assert_array_equal(got.view(np.uint8), exp.view(np.uint8))


@pytest.mark.parametrize("case", ["u3", "outlier", "pairs", "uniform", "clustered"])
def test_octree_stages(ctx, case):
    codes = TREE_CASES[case]
    if codes is None:
        codes = _unique_codes(120000, 33, case)
    c, tree, dev = _tree_on_gpu(ctx, codes)
    m = codes.size - 1
    e = oracle.edge_count(tree["prefix_n"], tree["parent"])
    p = oracle.partial_sum(e)
    p[0] = 0
    exp_nodes, k = oracle.octree(codes, tree, e, p)
    de = ctx.upload(e)
    dp = ctx.upload(p)
    nodes = ctx.empty(k * 15, np.int32)
    ctx.k_InitOctNodes(nodes, k, c, dev["prefix_n"], m, 0.0, 1024.0)
    ctx.k_MakeOctNodes(nodes, dp, de, c, dev["prefix_n"], dev["parent"], 0.0, 1024.0, m, k)
    ctx.k_LinkLeafNodes(nodes, dp, de, c, dev["has_leaf_left"], dev["has_leaf_right"],
                        dev["prefix_n"], dev["parent"], dev["left_child"], m, k)
    got = nodes.to_host().view(OCT_NODE_DTYPE)
    _nodes_equal(got, exp_nodes)

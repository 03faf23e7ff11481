"""The key->point permutation (not kept by the reference, which sorts bare
keys, kernels/cuda/02_sort.cu:566-574, and uniques in place,
kernels/03_unique.cpp:2-16): with it on, the sorted point indices equal
numpy's stable argsort of the Morton codes and each unique key's index is its
first occurrence; every other output is unchanged."""
import numpy as np
import pytest

import oracle
from pipeline_amd import clouds

pytestmark = pytest.mark.gpu

STAGES = ["sorted_keys", "unique_keys", "prefix_n", "has_leaf_left", "has_leaf_right",
          "left_child", "parent", "edge_count", "offsets"]


def _build(ctx, d, n, perm, mn=0.0, rg=1024.0):
    ctx.set_permutation(perm)
    try:
        ctx.build(d, n, mn, rg)
        return ctx.fetch()
    finally:
        ctx.set_permutation(False)


def _check_perm(out, s, codes):
    order = np.argsort(codes, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(out.sorted_index[s], order, err_msg="sorted_index")
    uk, first = np.unique(codes, return_index=True)
    np.testing.assert_array_equal(out.unique_keys[s], uk)
    np.testing.assert_array_equal(out.unique_index[s], first.astype(np.uint32),
                                  err_msg="unique_index")


@pytest.mark.parametrize("kind,n", [("uniform", 10_000_000), ("clustered", 1 << 24),
                                    ("clustered", 7681), ("uniform", 1)])
def test_plain_build_permutation(ctx, kind, n):
    gen = clouds.uniform if kind == "uniform" else clouds.clustered
    pts = gen(n, seed=114514)
    codes = oracle.morton(pts, nthreads=8)
    d = ctx.upload(pts)
    try:
        with_p = _build(ctx, d, n, True)
        _check_perm(with_p, 0, codes)
        without = _build(ctx, d, n, False)
        assert without.sorted_index is None
        for k in STAGES:
            np.testing.assert_array_equal(getattr(with_p, k)[0], getattr(without, k)[0], err_msg=k)
        np.testing.assert_array_equal(with_p.nodes[0].view(np.uint8),
                                      without.nodes[0].view(np.uint8))
    finally:
        d.free()


@pytest.mark.parametrize("graph", [False, True])
def test_batched_build_permutation(ctx, graph):
    pts, offs, smin, srg, _ = clouds.batch(30, 1, 400000, seed=17)
    d = ctx.upload(pts)
    ctx.set_permutation(True)
    ctx.set_graph_mode(graph)
    try:
        for _ in range(2 if graph else 1):
            ctx.build_batched(d, offs, smin, srg)
            out = ctx.fetch(offs)
            for s in range(offs.size - 1):
                seg = pts[offs[s]:offs[s + 1]]
                codes = oracle.morton(seg, float(smin[s]), float(srg[s]), nthreads=8)
                _check_perm(out, s, codes)
    finally:
        ctx.set_graph_mode(False)
        ctx.set_permutation(False)
        d.free()


@pytest.mark.parametrize("n", [2, 4095, 8193, 1_000_003])
def test_sort_pairs_and_unique_pairs(ctx, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 12, n, dtype=np.uint32) << 18  # heavy duplicates
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = ctx.upload(keys), ctx.upload(vals)
    ka, va = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32)
    try:
        ctx.k_SortPairs(dk, dv, ka, va, n)
        ctx.synchronize()
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(dk.to_host(), keys[order])
        np.testing.assert_array_equal(dv.to_host(), order.astype(np.uint32))
        u = ctx.k_CountUniquePairs(dk, dv, n)
        uk, first = np.unique(keys, return_index=True)
        assert u == uk.size
        np.testing.assert_array_equal(dk.to_host(u), uk)
        np.testing.assert_array_equal(dv.to_host(u), first.astype(np.uint32))
    finally:
        for a in (dk, dv, ka, va):
            a.free()

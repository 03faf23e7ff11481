"""ocp_k_SortKeys_static, the sort entry that allocates nothing, against
ocp_k_SortKeys (same keys, bit for bit): two calls in a row, the second on a
smaller input, each checked against the allocating entry and against numpy's
sort; sizes beyond the fixed scratch are refused."""
import numpy as np
import pytest

from pipeline_amd import STATIC_SORT_MAX_KEYS, OcpError

pytestmark = pytest.mark.gpu


def _sorted_by(ctx, fn, keys):
    dk = ctx.upload(keys)
    da = ctx.empty(keys.size, np.uint32)
    try:
        fn(dk, da, keys.size)
        ctx.synchronize()
        return dk.to_host()
    finally:
        dk.free()
        da.free()


def test_static_sort_matches_allocating_sort(ctx):
    rng = np.random.default_rng(2024)
    big = rng.integers(0, 1 << 30, 3_000_001, dtype=np.uint32)
    # a smaller second call with heavy duplicates and the top byte in use
    small = (rng.integers(0, 64, 123_457, dtype=np.uint32) << 26) | \
        rng.integers(0, 4, 123_457, dtype=np.uint32)
    for keys in (big, small):
        got = _sorted_by(ctx, ctx.k_SortKeys_static, keys)
        ref = _sorted_by(ctx, ctx.k_SortKeys, keys)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, np.sort(keys))


def test_static_sort_full_size_and_limit(ctx):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << 32, STATIC_SORT_MAX_KEYS, dtype=np.uint64).astype(np.uint32)
    got = _sorted_by(ctx, ctx.k_SortKeys_static, keys)
    np.testing.assert_array_equal(got, np.sort(keys))
    dk = ctx.empty(1, np.uint32)
    try:
        with pytest.raises(OcpError, match="at most 2"):
            ctx.k_SortKeys_static(dk, dk, STATIC_SORT_MAX_KEYS + 1)
    finally:
        dk.free()

"""Error path of the bounded lookback waits, forced from the host.

With a wait limit of 0 polls every decoupled-lookback wait (OneSweep,
unique, scan, learned-sort scatters) gives up at once; no wait can spin and
every kernel still runs to its end, with every access inside its buffers
(the first pass that gives up still sees keys that match its digit counts,
and every later lookback-stage kernel does nothing). The call that reports
the condition returns an error, a retried read of the dropped result fails
too, and the next call with the default limit matches the oracle. The
reference has no such path: its waits spin unbounded and its host code exits
on any error (include/cuda/common/helper_cuda.hpp:583-602).
"""
import numpy as np
import pytest

import oracle
from pipeline_amd import OcpError, clouds

pytestmark = pytest.mark.gpu

STAGES = ["prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
          "edge_count", "offsets"]


def _check(out, exp):
    np.testing.assert_array_equal(out.sorted_keys[0], exp["sorted"], err_msg="sorted")
    np.testing.assert_array_equal(out.unique_keys[0], exp["unique"], err_msg="unique")
    for k in STAGES:
        np.testing.assert_array_equal(getattr(out, k)[0], exp[k], err_msg=k)
    np.testing.assert_array_equal(out.nodes[0].view(np.uint8), exp["nodes"].view(np.uint8),
                                  err_msg="nodes")


@pytest.mark.parametrize("graph", [False, True])
def test_build_wait_gives_up_then_recovers(ctx, graph):
    pts = clouds.uniform(200000, seed=31)
    exp = oracle.build_plain(pts)
    d = ctx.upload(pts)
    ctx.set_graph_mode(graph)
    try:
        # a good build first, so that a graph (when on) is captured and replayed
        ctx.build(d, pts.shape[0], 0.0, 1024.0)
        _check(ctx.fetch(), exp)
        ctx.set_wait_limit(0)
        ctx.build(d, pts.shape[0], 0.0, 1024.0)
        with pytest.raises(OcpError, match="wait gave up"):
            ctx.results()
        with pytest.raises(OcpError, match="no build result"):
            ctx.results()
        with pytest.raises(OcpError, match="no build result"):
            ctx.counts(1)
        ctx.set_wait_limit(-1)
        ctx.build(d, pts.shape[0], 0.0, 1024.0)
        _check(ctx.fetch(), exp)
    finally:
        ctx.set_wait_limit(-1)
        ctx.set_graph_mode(False)
        d.free()


def test_stage_entries_report_at_synchronize(ctx):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 30, 300000, dtype=np.uint32)
    dk = ctx.upload(keys)
    da = ctx.empty(keys.size, np.uint32)
    vals = rng.integers(0, 4, 100000).astype(np.int32)
    dv = ctx.upload(vals)
    dp = ctx.empty(vals.size, np.int32)
    try:
        ctx.set_wait_limit(0)
        ctx.k_SortKeys(dk, da, keys.size)  # asynchronous: no error yet
        with pytest.raises(OcpError, match="wait gave up"):
            ctx.synchronize()
        ctx.synchronize()  # reported once, then cleared
        # a later entry that runs while the condition is pending does
        # nothing, and the condition is still reported once
        ctx.k_PartialSum(dv, 0, vals.size, dp)
        ctx.k_PartialSum(dv, 0, vals.size, dp)
        with pytest.raises(OcpError, match="wait gave up"):
            ctx.synchronize()
        ctx.set_wait_limit(-1)
        dk.copy_from(keys)
        ctx.k_SortKeys(dk, da, keys.size)
        ctx.synchronize()
        np.testing.assert_array_equal(dk.to_host(), np.sort(keys))
        ctx.k_PartialSum(dv, 0, vals.size, dp)
        ctx.synchronize()
        np.testing.assert_array_equal(dp.to_host(), np.cumsum(vals, dtype=np.int32))
        u = ctx.k_CountUnique(dk, keys.size)
        assert u == np.unique(keys).size
    finally:
        ctx.set_wait_limit(-1)
        for a in (dk, da, dv, dp):
            a.free()


def test_learned_sort_wait_gives_up_then_recovers(ctx):
    """The learned sort's scatters claim their runs without waiting; its
    waits are in the fallback's stable passes. A linear CDF over [0, 2^30)
    puts 2^20 keys spread over [0, 2^20) into 16 buckets of ~65K keys, each
    varying in its two low digits: two fallback passes, the second of which
    waits on earlier tiles of its bucket."""
    rng = np.random.default_rng(12)
    keys = rng.integers(0, 1 << 20, 1 << 20, dtype=np.uint32)
    dk = ctx.upload(keys)
    da = ctx.empty(keys.size, np.uint32)
    try:
        ctx.cdf_reset(0, (1 << 30) - 1)
        st = ctx.sort_learned(dk, da)
        np.testing.assert_array_equal(dk.to_host(), np.sort(keys))
        assert st["n_big"] == 16 and st["fallback_passes"] >= 2, st
        dk.copy_from(keys)
        ctx.set_wait_limit(0)
        with pytest.raises(OcpError, match="wait gave up"):
            ctx.sort_learned(dk, da)
        ctx.set_wait_limit(-1)
        dk.copy_from(keys)
        ctx.sort_learned(dk, da)
        np.testing.assert_array_equal(dk.to_host(), np.sort(keys))
    finally:
        ctx.set_wait_limit(-1)
        dk.free()
        da.free()

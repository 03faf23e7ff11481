"""The input side on the GPU: the reference's point generator
(kernels/cuda/00_init.cu:27-50) against its numpy restatement, and each
cloud's cube box reduced on the device against the host computation
(clouds.cube_box), with builds from device boxes equal to builds from host
boxes, byte for byte."""
import numpy as np
import pytest

import oracle
from pipeline_amd import clouds

pytestmark = pytest.mark.gpu

STAGES = ["sorted_keys", "unique_keys", "prefix_n", "has_leaf_left", "has_leaf_right",
          "left_child", "parent", "edge_count", "offsets"]


def _same_output(a, b, s):
    for k in STAGES:
        np.testing.assert_array_equal(getattr(a, k)[s], getattr(b, k)[s], err_msg=k)
    np.testing.assert_array_equal(a.nodes[s].view(np.uint8), b.nodes[s].view(np.uint8))


@pytest.mark.parametrize("n,seed", [(1, 114514), (1000, 114514), (1 << 20, 7), (3_000_001, 114514)])
def test_init_random_matches_restatement(ctx, n, seed):
    d = ctx.empty((n, 4), np.float32)
    try:
        ctx.k_InitRandomVec4(d, n, 0.0, 1024.0, seed)
        got = d.to_host()
        exp = oracle.gpu_random_vec4(n, 0.0, 1024.0, seed)
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
        assert got.min() >= 0.0 and got.max() < 1024.0
        if n >= 1 << 20:  # uniform per axis: each of 16 bins within 2% of n/16
            h = np.histogram(got[:, 0], bins=16, range=(0, 1024))[0]
            assert np.all(np.abs(h - n / 16) < 0.02 * n / 16)
    finally:
        d.free()


def test_init_random_feeds_the_build(ctx):
    n = 500_000
    d = ctx.empty((n, 4), np.float32)
    try:
        ctx.k_InitRandomVec4(d, n, 0.0, 1024.0, 114514)
        pts = d.to_host()
        ctx.build(d, n, 0.0, 1024.0)
        out = ctx.fetch()
        exp = oracle.build_plain(pts)
        np.testing.assert_array_equal(out.sorted_keys[0], exp["sorted"])
        np.testing.assert_array_equal(out.nodes[0].view(np.uint8), exp["nodes"].view(np.uint8))
    finally:
        d.free()


@pytest.mark.parametrize("float4", [True, False])
def test_device_boxes_match_host_boxes(ctx, float4):
    pts, offs, smin, srg, _ = clouds.batch(40, 1, 300000, seed=21, float4=float4)
    # an empty cloud and a one-point cloud among them
    offs = np.concatenate([[0, 0, 1], offs[1:] + 1]).astype(np.int32)
    one = np.zeros((1, pts.shape[1]), np.float32)
    one[0, :3] = (-3.5, 2.0, 7.25)
    pts = np.concatenate([one, pts]).astype(np.float32)
    smin = np.concatenate([[0.0, -3.5], smin]).astype(np.float32)
    srg = np.concatenate([[1.0, 1e-3], srg]).astype(np.float32)
    d = ctx.upload(pts)
    try:
        ctx.build_batched(d, offs, None, None)
        dev = ctx.fetch(offs)
        mn, rg = ctx.boxes()
        for s in range(offs.size - 1):
            seg = pts[offs[s]:offs[s + 1]]
            if seg.shape[0] == 0:
                assert (mn[s], rg[s]) == (0.0, 1.0)
                continue
            hm, hr = clouds.cube_box(seg)
            assert mn[s] == np.float32(hm) and rg[s] == np.float32(hr), s
        ctx.build_batched(d, offs, mn, rg)
        host = ctx.fetch(offs)
        for s in range(offs.size - 1):
            _same_output(dev, host, s)
    finally:
        d.free()


def test_autobox_plain_build(ctx):
    pts = clouds.clustered(300000, seed=3)
    pts[:, :3] = pts[:, :3] * np.float32(0.37) + np.float32(-120.0)
    mn, rg = clouds.cube_box(pts)
    d = ctx.upload(pts)
    try:
        ctx.build_autobox(d)
        got = ctx.fetch()
        bm, br = ctx.boxes()
        assert bm[0] == np.float32(mn) and br[0] == np.float32(rg)
        exp = oracle.build_plain(pts, mn, rg)
        np.testing.assert_array_equal(got.sorted_keys[0], exp["sorted"])
        np.testing.assert_array_equal(got.nodes[0].view(np.uint8), exp["nodes"].view(np.uint8))
    finally:
        d.free()


def test_device_boxes_512_clouds(ctx):
    pts, offs, smin, srg, _ = clouds.batch(512, seed=114514)
    d = ctx.upload(pts)
    try:
        ctx.build_batched(d, offs, None, None)
        mn, rg = ctx.boxes()
        np.testing.assert_array_equal(mn, smin)
        np.testing.assert_array_equal(rg, srg)
        u_dev, k_dev = ctx.counts(512)
        ctx.build_batched(d, offs, smin, srg)
        u_host, k_host = ctx.counts(512)
        np.testing.assert_array_equal(u_dev, u_host)
        np.testing.assert_array_equal(k_dev, k_host)
    finally:
        d.free()

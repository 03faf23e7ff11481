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


def _rotation(seed):
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(-0.5, 0.5)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)


@pytest.mark.parametrize("float4,n,in_place", [(True, 3_000_001, False), (False, 3_000_001, True),
                                               (True, 1000, True), (False, 0, False)])
def test_transform_points(ctx, float4, n, in_place):
    """ocp_transform_points against float64 A p + t (rounded once to float32
    the result differs by at most a few ulp of 1024), with clipping exact at
    both bounds and the fourth value of float4 points passed through bit for
    bit; the identity copies every point exactly."""
    pts = clouds.uniform(n, seed=3)
    host = pts if float4 else np.ascontiguousarray(pts[:, :3])
    rot = _rotation(11)
    c = np.array([512.0, 512.0, 512.0])
    aff = np.hstack([rot, (c + np.array([3.0, -40.0, 0.25]) - rot @ c)[:, None]])
    hi = float(np.nextafter(np.float32(1024.0), np.float32(0.0)))
    src = ctx.upload(host)
    dst = src if in_place else ctx.empty(host.shape, np.float32)
    try:
        ctx.transform_points(src, dst, aff, 0.0, hi)
        got = dst.to_host().reshape(host.shape)
        exp = np.clip(host[:, :3].astype(np.float64) @ rot.T + aff[:, 3], 0.0, hi)
        np.testing.assert_allclose(got[:, :3], exp, rtol=0, atol=1e-3)
        clipped_lo, clipped_hi = exp <= 0.0, exp >= hi
        assert np.all(got[:, :3][clipped_lo] == 0.0) and np.all(got[:, :3][clipped_hi] == np.float32(hi))
        if n >= 1000:
            assert clipped_lo.any()  # the -40 shift in y pushes points out at 0
        if float4:
            np.testing.assert_array_equal(got[:, 3].view(np.uint32), host[:, 3].view(np.uint32))
        back = ctx.upload(host)
        ctx.transform_points(back, back, np.hstack([np.eye(3), np.zeros((3, 1))]), 0.0, hi)
        np.testing.assert_array_equal(back.to_host().reshape(host.shape).view(np.uint32),
                                      host.view(np.uint32))
        back.free()
    finally:
        src.free()
        if not in_place:
            dst.free()


def test_transform_points_rejects_bad_arguments(ctx):
    from pipeline_amd import OcpError
    d = ctx.empty((10, 4), np.float32)
    try:
        with pytest.raises(OcpError):
            ctx.transform_points(d, d, np.zeros((3, 4)), 2.0, 1.0)
    finally:
        d.free()

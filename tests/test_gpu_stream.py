"""GPU tests of the streaming pipeline (ocp_ppl_*, the pipelined demo `ppl`
that apps/xmake.lua:7 and apps/README.md:5 name): every frame of a stream,
built while other frames are uploaded and built on other slots, equals the
oracle's build of that frame byte for byte."""
import numpy as np
import pytest

import oracle
from pipeline_amd import OcpError, Pipeline, TreeReport, clouds

pytestmark = pytest.mark.gpu


def _frames(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(count):
        n = int(rng.integers(lo, hi))
        gen = clouds.uniform if f % 2 == 0 else clouds.clustered
        out.append(gen(n, seed=seed * 100 + f))
    return out


def _check_frame(ctx, pts):
    exp = oracle.build_plain(pts)
    out = ctx.fetch()
    np.testing.assert_array_equal(out.sorted_keys[0], exp["sorted"])
    np.testing.assert_array_equal(out.unique_keys[0], exp["unique"])
    for k in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
              "edge_count", "offsets"):
        np.testing.assert_array_equal(getattr(out, k)[0], exp[k], err_msg=k)
    np.testing.assert_array_equal(out.nodes[0].view(np.uint8), exp["nodes"].view(np.uint8))


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_stream_device_frames(depth):
    frames = _frames(7, 20000, 400000, seed=depth)
    with Pipeline(depth) as ppl:
        dev = [ppl.context(0).upload(p) for p in frames]
        for f in range(min(depth, len(frames))):
            ppl.submit(f % depth, dev[f], frames[f].shape[0])
        for f in range(len(frames)):
            slot = f % depth
            ppl.wait(slot)
            _check_frame(ppl.context(slot), frames[f])
            if f + depth < len(frames):
                g = f + depth
                ppl.submit(g % depth, dev[g], frames[g].shape[0])


@pytest.mark.parametrize("depth", [2, 3])
def test_stream_staged_host_frames(depth):
    """The producer writes each frame into the slot's pinned buffer; the
    upload runs on the slot's stream beside the other slots' builds."""
    frames = _frames(8, 1000, 300000, seed=10 + depth)
    cap = max(p.shape[0] for p in frames)
    with Pipeline(depth, max_points=cap) as ppl:
        for f in range(len(frames) + depth):
            slot = f % depth
            if f >= depth:
                ppl.wait(slot)
                _check_frame(ppl.context(slot), frames[f - depth])
            if f < len(frames):
                n = frames[f].shape[0]
                buf = ppl.host_buffer(slot, n)
                buf[:] = frames[f]
                ppl.submit(slot, None, n)


def test_stream_large_frames_overlap():
    """Twelve 2M-point frames over three slots, device-resident: the
    overlapped builds stay exact and the validator finds nothing."""
    frames = [clouds.uniform(2_000_000, seed=500 + f) if f % 3 else
              clouds.clustered(2_000_000, seed=500 + f) for f in range(12)]
    with Pipeline(3) as ppl:
        dev = [ppl.context(0).upload(p) for p in frames]
        for f in range(3):
            ppl.submit(f, dev[f], 2_000_000)
        for f in range(12):
            slot = f % 3
            ppl.wait(slot)
            ctx = ppl.context(slot)
            _check_frame(ctx, frames[f])
            rep = ctx.validate_octree()
            assert all(rep[k] == 0 for k in TreeReport.VIOLATIONS), rep
            if f + 3 < 12:
                ppl.submit(slot, dev[f + 3], 2_000_000)


def test_stream_argument_errors():
    with Pipeline(2, max_points=1000) as ppl:
        with pytest.raises(OcpError, match="slot out of range"):
            ppl.submit(2, None, 10)
        with pytest.raises(OcpError, match="max_points"):
            ppl.submit(0, None, 1001)
        buf = ppl.host_buffer(1, 1000)
        buf[:] = clouds.uniform(1000, seed=3)
        ppl.submit(1, None, 1000)
        ppl.wait(1)
        _check_frame(ppl.context(1), clouds.uniform(1000, seed=3))

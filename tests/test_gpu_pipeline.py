"""GPU parity of the fused device-resident build (plain, graph replay,
batched) against the oracle's full chain, stage output by stage output."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from octree_checks import check_octree
from pipeline_amd import TreeReport, clouds

pytestmark = pytest.mark.gpu

STAGES = ["prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
          "edge_count", "offsets"]


def _check_segment(out, s, exp):
    np.testing.assert_array_equal(out.sorted_keys[s], exp["sorted"], err_msg="sorted")
    assert int(out.n_unique[s]) == exp["n_unique"]
    np.testing.assert_array_equal(out.unique_keys[s], exp["unique"], err_msg="unique")
    for k in STAGES:
        np.testing.assert_array_equal(getattr(out, k)[s], exp[k], err_msg=k)
    assert int(out.n_oct[s]) == exp["n_oct"]
    np.testing.assert_array_equal(out.nodes[s].view(np.uint8), exp["nodes"].view(np.uint8),
                                  err_msg="nodes")


def _plain(ctx, pts, mn=0.0, rg=1024.0):
    d = ctx.upload(pts)
    ctx.build(d, pts.shape[0], mn, rg)
    out = ctx.fetch()
    d.free()
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 17, 4096, 7681, 307200])
def test_plain_uniform(ctx, n):
    pts = clouds.uniform(n, seed=114514 + n)
    out = _plain(ctx, pts)
    _check_segment(out, 0, oracle.build_plain(pts))


@pytest.mark.parametrize("n", [5000, 262144])
def test_plain_clustered(ctx, n):
    pts = clouds.clustered(n, seed=7 + n)
    out = _plain(ctx, pts)
    _check_segment(out, 0, oracle.build_plain(pts))


def test_plain_empty_between_builds(ctx):
    """An empty cloud between two builds: no unique key and no node, and the
    next build is exact (the node offsets of the empty build come from the
    offsets kernel, those of the others from the edge scan's last group)."""
    pts = clouds.uniform(5000, seed=3)
    d = ctx.upload(pts)
    try:
        for n in (5000, 0, 1, 5000):
            ctx.build(d, n, 0.0, 1024.0)
            u, k = ctx.counts(1)
            if n == 0:
                assert int(u[0]) == 0 and int(k[0]) == 0, (u, k)
            else:
                _check_segment(ctx.fetch(), 0, oracle.build_plain(pts[:n]))
    finally:
        d.free()


def test_plain_all_duplicates_and_float3(ctx):
    pts = np.tile(np.array([[1.5, 2.5, 3.5, 0.0]], np.float32), (10000, 1))
    _check_segment(_plain(ctx, pts), 0, oracle.build_plain(pts))
    p3 = clouds.uniform(70001, seed=3, float4=False)
    _check_segment(_plain(ctx, p3), 0, oracle.build_plain(p3))
    pts = clouds.clustered(30000, seed=4)
    pts[:, :3] = pts[:, :3] * np.float32(0.25) + np.float32(-40.0)
    _check_segment(_plain(ctx, pts, -40.0, 256.0), 0, oracle.build_plain(pts, -40.0, 256.0))


def test_graph_replay_matches_eager(ctx):
    pts = clouds.clustered(200000, seed=99)
    exp = oracle.build_plain(pts)
    d = ctx.upload(pts)
    ctx.set_graph_mode(True)
    try:
        for _ in range(3):
            ctx.build(d, pts.shape[0], 0.0, 1024.0)
            ctx.synchronize()
            _check_segment(ctx.fetch(), 0, exp)
        # new data in the same buffer: the replay must see it
        pts2 = clouds.uniform(200000, seed=100)
        d.copy_from(pts2)
        ctx.build(d, pts2.shape[0], 0.0, 1024.0)
        _check_segment(ctx.fetch(), 0, oracle.build_plain(pts2))
    finally:
        ctx.set_graph_mode(False)
        d.free()


def _batched_case(n_clouds, lo, hi, seed):
    return clouds.batch(n_clouds, lo, hi, seed=seed)


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_per_cloud_oracle(ctx, graph):
    pts, offs, smin, srg, _ = _batched_case(24, 1, 60000, seed=5)
    # include empty and single-point clouds
    offs = np.concatenate([[0, 0], offs[1:] + 0]).astype(np.int32)
    smin = np.concatenate([[0.0], smin]).astype(np.float32)
    srg = np.concatenate([[1.0], srg]).astype(np.float32)
    d = ctx.upload(pts)
    ctx.set_graph_mode(graph)
    try:
        for rep in range(2 if graph else 1):
            ctx.build_batched(d, offs, smin, srg)
            out = ctx.fetch(offs)
            for s in range(offs.size - 1):
                seg = pts[offs[s]:offs[s + 1]]
                if seg.shape[0] == 0:
                    assert out.n_unique[s] == 0 and out.n_oct[s] == 0
                    continue
                _check_segment(out, s, oracle.build_plain(seg, float(smin[s]), float(srg[s])))
    finally:
        ctx.set_graph_mode(False)
        d.free()


def _segment_bytes(out, s):
    return [out.sorted_keys[s], out.unique_keys[s]] + [getattr(out, k)[s] for k in STAGES] + \
        [out.nodes[s].view(np.uint8), out.n_unique[s:s + 1], out.n_oct[s:s + 1]]


@pytest.mark.timeout(900)
def test_batched_512_clouds_full_workload(ctx):
    """The batched workload at its real size: 512 clouds of 4,096 to ~2M
    points (log-uniform, uniform or clustered, each with its own box), one
    segmented launch per stage. Every cloud above 262,144 points (64 scan
    tiles, where the two-level lookback of a later segment starts) and 64
    others are compared byte for byte with the oracle; a graph replay must
    give the eager build's bytes for every cloud."""
    pts, offs, smin, srg, _ = clouds.batch(512, seed=114514)
    sizes = np.diff(offs)
    d = ctx.upload(pts)
    try:
        ctx.set_graph_mode(False)
        ctx.build_batched(d, offs, smin, srg)
        eager = ctx.fetch(offs)
        big = np.nonzero(sizes > 262144)[0]
        rest = np.nonzero(sizes <= 262144)[0]
        picked = np.concatenate([big, rest[np.linspace(0, rest.size - 1, 64).astype(int)]])
        assert big.size > 100 and picked.size == big.size + 64

        def one(s):
            seg = pts[offs[s]:offs[s + 1]]
            _check_segment(eager, s, oracle.build_plain(seg, float(smin[s]), float(srg[s])))
            return s

        with ThreadPoolExecutor(max_workers=8) as ex:
            done = list(ex.map(one, picked.tolist()))
        assert len(done) == picked.size
        ctx.set_graph_mode(True)
        for _ in range(2):  # capture, then replay
            ctx.build_batched(d, offs, smin, srg)
            rep = ctx.fetch(offs)
            for s in range(512):
                for a, b in zip(_segment_bytes(rep, s), _segment_bytes(eager, s)):
                    assert np.array_equal(a, b), f"cloud {s}"
    finally:
        ctx.set_graph_mode(False)
        d.free()


def _device_checks(ctx, out, written):
    """The device validator agrees with the CPU structure check, and the
    compact form expands back to the same node bytes."""
    rep = ctx.validate_octree()
    assert all(rep[f] == 0 for f in TreeReport.VIOLATIONS), rep
    assert rep["written"] == written
    cn, ch = ctx.compact_octree()
    r = ctx.results()
    k = int(out.n_oct[0])
    back = ctx.expand_octree(cn, ch, r.oct_start, r.n_seg, None, None, 0.0, 1024.0)
    np.testing.assert_array_equal(back.to_host(k * 60), out.nodes[0].view(np.uint8))
    for a in (cn, ch, back):
        a.free()


def test_plain_10M_uniform_full_oracle(ctx):
    """The plain GPU workload at its real size: 10M uniform float4 points,
    seed 114514, every stage output against the oracle, plus the octree's
    order-independent structure (written slots, zero gaps, grid-aligned
    corners, power-of-two cells) checked without the oracle."""
    pts = clouds.uniform(10_000_000, seed=114514)
    exp = oracle.build_plain(pts)
    out = _plain(ctx, pts)
    written = check_octree(out.nodes[0], out.unique_keys[0], out.prefix_n[0],
                           out.edge_count[0], out.offsets[0])
    _check_segment(out, 0, exp)
    _device_checks(ctx, out, written)


def test_plain_2p24_clustered_full_oracle(ctx):
    """The clustered workload at its real size (2^24 points, heavy duplicate
    keys), every stage output against the oracle."""
    pts = clouds.clustered(1 << 24, seed=114514)
    exp = oracle.build_plain(pts)
    assert exp["n_unique"] < (1 << 24) // 2  # the duplicate-heavy case
    out = _plain(ctx, pts)
    written = check_octree(out.nodes[0], out.unique_keys[0], out.prefix_n[0],
                           out.edge_count[0], out.offsets[0])
    _check_segment(out, 0, exp)
    _device_checks(ctx, out, written)


def test_tuned_configs_give_identical_outputs(ctx):
    """Every tile configuration the autotuner can pick gives the same bytes
    (plain and batched)."""
    pts = clouds.clustered(300000, seed=77)
    exp = oracle.build_plain(pts)
    bpts, offs, smin, srg, _ = _batched_case(6, 1000, 80000, seed=8)
    d = ctx.upload(pts)
    db = ctx.upload(bpts)
    base = ctx.get_tune()
    configs = ([dict(sort_waves=w, sort_kpt=k) for w, k in
                ((8, 15), (4, 16), (4, 24), (8, 8), (8, 12), (8, 16), (16, 8))] +
               [dict(sort_bits=10, sort10_kpt=k, sort_msd=m) for k in (8, 16) for m in (0, 1)] +
               [dict(scan_ipt=i) for i in (8, 16, 32)] +
               [dict(morton_ppt=p, morton_grid=g) for p in (4, 8, 16, 32)
                for g in (256, 512, 2048)] +
               [dict(oct_order=1)])
    try:
        for cfg in configs:
            ctx.set_tune(**{**base, **cfg})
            ctx.build(d, pts.shape[0], 0.0, 1024.0)
            _check_segment(ctx.fetch(), 0, exp)
            ctx.build_batched(db, offs, smin, srg)
            out = ctx.fetch(offs)
            for s in (0, offs.size - 2):
                seg = bpts[offs[s]:offs[s + 1]]
                _check_segment(out, s, oracle.build_plain(seg, float(smin[s]), float(srg[s])))
    finally:
        ctx.set_tune(**base)
        d.free()
        db.free()


def test_octree_group_order_full_size(ctx):
    """The octree's other work order (a tile's chains grouped by the size of
    their group walk) writes the same node bytes as index order on both GPU
    workloads at full size and on a batch built from a graph."""
    base = ctx.get_tune()
    cases = [("uniform 10M", clouds.uniform(10_000_000, seed=114514)),
             ("clustered 2^24", clouds.clustered(1 << 24, seed=114514))]
    try:
        for name, pts in cases:
            print(f"case {name}", flush=True)
            d = ctx.upload(pts)
            got = {}
            for order in (0, 1):
                ctx.set_tune(**{**base, "oct_order": order})
                ctx.build(d, pts.shape[0], 0.0, 1024.0)
                got[order] = ctx.fetch().nodes[0].view(np.uint8)
            assert got[0].size > 0
            np.testing.assert_array_equal(got[1], got[0], err_msg=name)
            d.free()
        bpts, offs, smin, srg, _ = _batched_case(40, 1000, 300000, seed=9)
        db = ctx.upload(bpts)
        got = {}
        for order in (0, 1):
            ctx.set_tune(**{**base, "oct_order": order})
            ctx.set_graph_mode(True)
            ctx.build_batched(db, offs, smin, srg)
            ctx.build_batched(db, offs, smin, srg)
            out = ctx.fetch(offs)
            got[order] = [out.nodes[k].view(np.uint8).copy() for k in range(offs.size - 1)]
        for k in range(offs.size - 1):
            np.testing.assert_array_equal(got[1][k], got[0][k], err_msg=f"cloud {k}")
        db.free()
    finally:
        ctx.set_graph_mode(False)
        ctx.set_tune(**base)


def _msd_fallback_ran(ctx):
    names = list(ctx.kernel_ms())
    assert "msd_scatter" in names and "msd_buckets" in names
    assert ("msd_lsd_fallback" in names) != ("msd_lsd_skipped" in names)
    return "msd_lsd_fallback" in names


def test_msd_sort_full_size(ctx):
    """The MSD form of the sort (one 10-bit scatter by the top digit, then
    every bucket sorted in LDS by one workgroup) gives the oracle's bytes on
    10M uniform points, where every bucket fits in LDS and the LSD passes
    are skipped, and on 2^24 clustered points, whose large buckets send the
    build through the three LSD passes; then on a batch that holds both
    kinds of cloud (eager and from a graph, twice each) and on a batch of
    uniform clouds, all of whose buckets fit; builds with the permutation
    keep stable 8-bit passes."""
    base = ctx.get_tune()
    try:
        ctx.set_tune(**{**base, "sort_bits": 10, "sort_msd": 1})
        ctx.set_stage_timing(True)
        for name, pts, fallback in (
                ("uniform 10M", clouds.uniform(10_000_000, seed=114514), False),
                ("clustered 2^24", clouds.clustered(1 << 24, seed=114514), True)):
            print(f"case {name}", flush=True)
            d = ctx.upload(pts)
            ctx.build(d, pts.shape[0], 0.0, 1024.0)
            assert _msd_fallback_ran(ctx) == fallback, name
            _check_segment(ctx.fetch(), 0, oracle.build_plain(pts))
            d.free()
        sizes = [1, 777, 65536, 300_000, 2_000_000]
        upts = np.concatenate([clouds.uniform(n, seed=40 + k) for k, n in enumerate(sizes)])
        uoffs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        ubox = (np.zeros(len(sizes), np.float32), np.full(len(sizes), 1024.0, np.float32))
        for name, (bpts, offs, smin, srg), fallback in (
                ("mixed batch", _batched_case(10, 1000, 300_000, seed=31)[:4], None),
                ("uniform batch", (upts, uoffs) + ubox, False)):
            print(f"case {name}", flush=True)
            exps = [oracle.build_plain(bpts[offs[s]:offs[s + 1]], float(smin[s]), float(srg[s]))
                    for s in range(offs.size - 1)]
            db = ctx.upload(bpts)
            for graph in (False, True):
                ctx.set_graph_mode(graph)
                for _ in range(2):
                    ctx.build_batched(db, offs, smin, srg)
                    if not graph and fallback is not None:
                        assert _msd_fallback_ran(ctx) == fallback, name
                    out = ctx.fetch(offs)
                    for s in range(offs.size - 1):
                        _check_segment(out, s, exps[s])
            ctx.set_graph_mode(False)
            db.free()
        ctx.set_stage_timing(False)
        small = clouds.uniform(300_000, seed=5)
        ds = ctx.upload(small)
        ctx.set_permutation(True)
        ctx.build(ds, small.shape[0], 0.0, 1024.0)
        out = ctx.fetch()
        keys = oracle.morton(small, 0.0, 1024.0)
        np.testing.assert_array_equal(out.sorted_index[0], np.argsort(keys, kind="stable"))
        _check_segment(out, 0, oracle.build_plain(small))
        ds.free()
    finally:
        ctx.set_stage_timing(False)
        ctx.set_permutation(False)
        ctx.set_graph_mode(False)
        ctx.set_tune(**base)


def test_msd_sort_small_and_many_clouds(ctx):
    """The MSD form on tiny clouds (buckets of zero to a few keys, every key
    equal, a cloud of one point), on a 17-point cloud built from a graph,
    and on a batch of more than 2048 clouds, which keeps the LSD passes."""
    base = ctx.get_tune()
    try:
        ctx.set_tune(**{**base, "sort_bits": 10, "sort_msd": 1})
        for n in (1, 2, 3, 17, 4096, 70001):
            _check_segment(_plain(ctx, clouds.uniform(n, seed=60 + n)), 0,
                           oracle.build_plain(clouds.uniform(n, seed=60 + n)))
        same = np.tile(np.array([[7.5, 1.5, 900.25, 0.0]], np.float32), (5000, 1))
        _check_segment(_plain(ctx, same), 0, oracle.build_plain(same))
        pts = clouds.uniform(17, seed=9)
        d = ctx.upload(pts)
        ctx.set_graph_mode(True)
        for _ in range(3):
            ctx.build(d, 17, 0.0, 1024.0)
            _check_segment(ctx.fetch(), 0, oracle.build_plain(pts))
        ctx.set_graph_mode(False)
        d.free()
        bpts, offs, smin, srg, _ = _batched_case(2100, 1, 400, seed=61)
        db = ctx.upload(bpts)
        ctx.set_stage_timing(True)
        ctx.build_batched(db, offs, smin, srg)
        names = list(ctx.kernel_ms())
        assert "onesweep_pass0" in names and "msd_scatter" not in names
        ctx.set_stage_timing(False)
        out = ctx.fetch(offs)
        for s in list(range(0, 2100, 97)) + [2099]:
            seg = bpts[offs[s]:offs[s + 1]]
            _check_segment(out, s, oracle.build_plain(seg, float(smin[s]), float(srg[s])))
        db.free()
    finally:
        ctx.set_stage_timing(False)
        ctx.set_graph_mode(False)
        ctx.set_tune(**base)


@pytest.mark.parametrize("kpt", [8, 16])
def test_ten_bit_digits_full_size(ctx, kpt):
    """Three passes of 10-bit digits (the autotuner's other pass count) give
    the oracle's bytes on the 10M workload, on a batch whose clouds span many
    tiles (eager and from a graph), and leave builds with the permutation on
    four stable 8-bit passes."""
    base = ctx.get_tune()
    pts = clouds.uniform(10_000_000, seed=114514)
    bpts, offs, smin, srg, _ = _batched_case(12, 4096, 2_000_000, seed=21)
    d = ctx.upload(pts)
    db = ctx.upload(bpts)
    try:
        ctx.set_tune(**{**base, "sort_bits": 10, "sort10_kpt": kpt})
        ctx.build(d, pts.shape[0], 0.0, 1024.0)
        _check_segment(ctx.fetch(), 0, oracle.build_plain(pts))
        exps = [oracle.build_plain(bpts[offs[s]:offs[s + 1]], float(smin[s]), float(srg[s]))
                for s in range(offs.size - 1)]
        for graph in (False, True):
            ctx.set_graph_mode(graph)
            for _ in range(2):
                ctx.build_batched(db, offs, smin, srg)
                out = ctx.fetch(offs)
                for s in range(offs.size - 1):
                    _check_segment(out, s, exps[s])
        ctx.set_graph_mode(False)
        small = pts[:300000]
        ds = ctx.upload(small)
        ctx.set_permutation(True)
        ctx.build(ds, small.shape[0], 0.0, 1024.0)
        out = ctx.fetch()
        keys = oracle.morton(small, 0.0, 1024.0)
        np.testing.assert_array_equal(out.sorted_index[0], np.argsort(keys, kind="stable"))
        _check_segment(out, 0, oracle.build_plain(small))
    finally:
        ctx.set_permutation(False)
        ctx.set_graph_mode(False)
        ctx.set_tune(**base)
        d.free()
        db.free()


def test_autotune_cache(ctx, tmp_path):
    """The tuning cache: a first call tunes (pass count included) and appends
    the workload class; a second call applies it without timing; the build
    stays exact under the chosen configuration."""
    pts = clouds.uniform(200_000, seed=3)
    d = ctx.upload(pts)
    base = ctx.get_tune()
    path = tmp_path / "tune.txt"
    try:
        t1 = ctx.autotune(d, reps=1, cache=str(path))
        assert not ctx.tune_from_cache and t1["sort_bits"] in (8, 10)
        ctx.set_tune(**base)
        t2 = ctx.autotune(d, reps=1, cache=str(path))
        assert ctx.tune_from_cache and t2 == t1 and ctx.get_tune() == t1
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and "/plain/f4/n2^18/" in lines[0]
        assert len(lines[0].split()) == 10
        ctx.build(d, pts.shape[0], 0.0, 1024.0)
        _check_segment(ctx.fetch(), 0, oracle.build_plain(pts))
        # Lines from before the octree order and the MSD form were tuned
        # still load, with those fields at 0; the last line of a class wins.
        key = lines[0].split()[0]
        for line, want in ((f"{key} 8 16 16 1024 16 10 16", {"oct_order": 0, "sort_msd": 0}),
                           (f"{key} 8 16 16 1024 16 10 16 1", {"oct_order": 1, "sort_msd": 0}),
                           (f"{key} 8 16 16 512 8 10 16 0 1", {"oct_order": 0, "sort_msd": 1})):
            path.write_text(lines[0] + "\n" + line + "\n")
            ctx.set_tune(**base)
            t = ctx.autotune(d, reps=1, cache=str(path))
            assert ctx.tune_from_cache and t["sort_bits"] == 10, line
            assert {k: t[k] for k in want} == want, line
            ctx.build(d, pts.shape[0], 0.0, 1024.0)
            _check_segment(ctx.fetch(), 0, oracle.build_plain(pts))
    finally:
        ctx.set_tune(**base)
        d.free()


def test_workspace_size_matches_device_memory():
    """ocp_workspace_size predicts the device memory a fresh context takes
    for its first build (measured as the drop in free device memory)."""
    from pipeline_amd import Context, workspace_size
    n = 10_000_000
    pts = clouds.uniform(n, seed=5)
    c = Context(0)
    try:
        d = c.upload(pts)
        c.synchronize()
        free0, _ = c.device_memory()
        c.build(d, n, 0.0, 1024.0)
        c.synchronize()
        free1, _ = c.device_memory()
        used = free0 - free1
        want = workspace_size(n)["build_bytes"]
        assert abs(used - want) <= 64 * 2**20 + 0.02 * want, (used, want)
    finally:
        c.close()

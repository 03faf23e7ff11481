"""GPU tests of the octree's structural validator (ocp_validate_octree, in the
role of checkTree, apps/cpu_demo/main.cpp:13-35) and the compact node format
(ocp_compact_octree / ocp_expand_octree).

The validator is pinned against the CPU structure check (octree_checks.py):
it reports no violation on outputs that check passes, the same number of
written slots, and exactly one violation of the planted kind for each
mistake planted into the node array. The compact form is pinned by its exact
round trip to the 60-byte nodes and by decoding its fields on the CPU (the
corner is the Morton decode of the prefix, morton.h:45-80, the cell size
range / 2^depth)."""
import ctypes

import numpy as np
import pytest

from octree_checks import check_octree
from pipeline_amd import OCT_NODE_DTYPE, TreeReport, clouds, d2h

pytestmark = pytest.mark.gpu


def _third_bits(m):
    x = m.astype(np.uint32) & np.uint32(0x9249249)
    x = (x ^ (x >> np.uint32(2))) & np.uint32(0x30c30c3)
    x = (x ^ (x >> np.uint32(4))) & np.uint32(0x0300f00f)
    x = (x ^ (x >> np.uint32(8))) & np.uint32(0x30000ff)
    x = (x ^ (x >> np.uint32(16))) & np.uint32(0x3ff)
    return x


def _links_in_octant(nodes):
    """(links, in_octant) over node-mask links, float32 as on the device."""
    links = inside = 0
    k = nodes.size
    for g in np.nonzero(nodes["child_node_mask"])[0]:
        nd = nodes[g]
        h = np.float32(nd["cell_size"] * np.float32(0.5))
        for c in range(8):
            if not (int(nd["child_node_mask"]) >> c) & 1:
                continue
            links += 1
            t = int(nd["children"][c])
            if not 0 <= t < k:
                continue
            ch = nodes[t]
            lo = [np.float32(nd["corner"][a] + (h if (c >> a) & 1 else np.float32(0)))
                  for a in range(3)]
            cs = ch["cell_size"]
            ok = cs > 0 and cs <= h
            for a in range(3):
                ok = ok and ch["corner"][a] >= lo[a] and \
                    np.float32(ch["corner"][a] + cs) <= np.float32(lo[a] + h)
            inside += bool(ok)
    return links, inside


def _check_compact(cn, ch, nodes, root_level, mn, rg):
    """CPU decode of one segment's compact form against its 60-byte nodes."""
    raw = nodes.view(np.uint8).reshape(-1, 60)
    written = raw.any(axis=1)
    assert np.all((cn["level"] == 0xff) == ~written)
    # child_set and the children array
    nz = nodes["children"] != 0
    np.testing.assert_array_equal(
        cn["child_set"], (nz * (1 << np.arange(8))).sum(axis=1).astype(np.uint8))
    np.testing.assert_array_equal(ch, nodes["children"][nz])
    first = np.concatenate([[0], np.cumsum(nz.sum(axis=1))[:-1]])
    np.testing.assert_array_equal(cn["first_child"] - cn["first_child"][0], first)
    w = written
    level = (cn["level"][w] & 15).astype(np.int64)
    depth = (cn["level"][w] >> 4).astype(np.int64)
    np.testing.assert_array_equal(level, root_level + depth)
    np.testing.assert_array_equal(
        nodes["cell_size"][w],
        (np.float32(rg) / (2.0 ** depth).astype(np.float32)).astype(np.float32))
    shift = np.where(level >= 10, 0, 30 - 3 * level).astype(np.uint32)
    pre = cn["prefix"][w]
    assert np.all(((pre >> shift) << shift) == pre), "prefix off its level"
    for a in range(3):
        t = _third_bits(pre >> np.uint32(a)).astype(np.float32) / np.float32(1024.0)
        np.testing.assert_array_equal(nodes["corner"][w, a],
                                      (t * np.float32(rg)).astype(np.float32) + np.float32(mn))
    np.testing.assert_array_equal(cn["node_mask"][w], nodes["child_node_mask"][w])
    np.testing.assert_array_equal(cn["leaf_mask"][w], nodes["child_leaf_mask"][w])


def _round_trip(ctx, r, n_nodes):
    cn, ch = ctx.compact_octree()
    assert cn.size == n_nodes
    back = ctx.expand_octree(cn, ch, r.oct_start, r.n_seg, r.seg_min or None,
                             r.seg_range or None, 0.0, 1024.0)
    got = back.to_host(n_nodes * 60)
    want = d2h(ctx, r.nodes, n_nodes * 60, np.uint8)
    np.testing.assert_array_equal(got, want)
    out = cn.to_host(), ch.to_host()
    for a in (cn, ch, back):
        a.free()
    return out


@pytest.mark.parametrize("kind,n", [("uniform", 1), ("uniform", 2), ("uniform", 3000),
                                    ("uniform", 307200), ("clustered", 200000)])
def test_validate_and_compact_plain(ctx, kind, n):
    pts = getattr(clouds, kind)(n, seed=31 + n)
    d = ctx.upload(pts)
    ctx.build(d, n, 0.0, 1024.0)
    out = ctx.fetch()
    d.free()
    nodes = out.nodes[0]
    rep = ctx.validate_octree()
    written = check_octree(nodes, out.unique_keys[0], out.prefix_n[0], out.edge_count[0],
                           out.offsets[0])
    assert all(rep[f] == 0 for f in TreeReport.VIOLATIONS), rep
    assert rep["slots"] == nodes.size and rep["written"] == written
    if n <= 3000:
        assert (rep["links"], rep["links_in_octant"]) == _links_in_octant(nodes)
    r = ctx.results()
    cn, ch = _round_trip(ctx, r, nodes.size)
    root_level = int(out.prefix_n[0][0]) // 3 if out.n_unique[0] >= 2 else 0
    _check_compact(cn, ch, nodes, root_level, 0.0, 1024.0)
    assert cn.nbytes + ch.nbytes < nodes.nbytes / 2


def test_validate_and_compact_batched(ctx):
    pts, offs, smin, srg, _ = clouds.batch(40, 1, 60000, seed=5)
    # an empty cloud and a one-point cloud ahead of them
    pts = np.concatenate([pts[:1], pts])
    offs = np.concatenate([[0, 0], np.asarray(offs, np.int64) + 1]).astype(np.int32)
    smin = np.concatenate([[0.0], smin[:1], smin]).astype(np.float32)
    srg = np.concatenate([[1.0], srg[:1], srg]).astype(np.float32)
    d = ctx.upload(pts)
    ctx.build_batched(d, offs, smin, srg)
    out = ctx.fetch(offs)
    d.free()
    rep = ctx.validate_octree()
    assert all(rep[f] == 0 for f in TreeReport.VIOLATIONS), rep
    written = 0
    for s in range(offs.size - 1):
        if out.n_oct[s] == 0:
            continue
        written += check_octree(out.nodes[s], out.unique_keys[s], out.prefix_n[s],
                                out.edge_count[s], out.offsets[s], smin[s], srg[s])
    assert rep["written"] == written
    r = ctx.results()
    total = int(sum(int(k) for k in out.n_oct))
    cn, ch = _round_trip(ctx, r, total)
    ost = d2h(ctx, r.oct_start, r.n_seg + 1, np.int32)
    for s in range(offs.size - 1):
        a, b = int(ost[s]), int(ost[s + 1])
        if a == b:
            continue
        c0 = int(cn["first_child"][a])
        c1 = c0 + int(sum(bin(int(x)).count("1") for x in cn["child_set"][a:b]))
        rl = int(out.prefix_n[s][0]) // 3 if out.n_unique[s] >= 2 else 0
        _check_compact(cn[a:b], ch[c0:c1], out.nodes[s], rl, smin[s], srg[s])


def test_validate_autobox_uses_device_boxes(ctx):
    pts = clouds.clustered(100000, seed=12)
    pts[:, :3] = pts[:, :3] * np.float32(0.01) + np.float32(-3.0)
    d = ctx.upload(pts)
    ctx.build_autobox(d)
    mn, rg = ctx.boxes()
    out = ctx.fetch()
    d.free()
    rep = ctx.validate_octree()
    assert all(rep[f] == 0 for f in TreeReport.VIOLATIONS), rep
    assert rep["written"] == check_octree(out.nodes[0], out.unique_keys[0], out.prefix_n[0],
                                          out.edge_count[0], out.offsets[0], mn[0], rg[0])
    r = ctx.results()
    cn, ch = _round_trip(ctx, r, out.nodes[0].size)
    rl = int(out.prefix_n[0][0]) // 3
    _check_compact(cn, ch, out.nodes[0], rl, mn[0], rg[0])


def _write_node(ctx, r, slot, node):
    buf = np.ascontiguousarray(np.array([node], OCT_NODE_DTYPE))
    ctx._check(ctx.lib.ocp_memcpy_h2d(ctx.h, ctypes.c_void_p(r.nodes + 60 * slot),
                                      buf.ctypes.data_as(ctypes.c_void_p), 60))


def test_validator_counts_planted_mistakes(ctx):
    """Each mistake, written into the node array of a correct build, shows up
    as exactly one violation of its kind (the CPU check rejects it too)."""
    pts = clouds.uniform(50000, seed=77)
    d = ctx.upload(pts)
    ctx.build(d, 50000, 0.0, 1024.0)
    out = ctx.fetch()
    d.free()
    r = ctx.results()
    nodes = out.nodes[0].copy()
    args = (out.unique_keys[0], out.prefix_n[0], out.edge_count[0], out.offsets[0])
    base = ctx.validate_octree()
    assert all(base[f] == 0 for f in TreeReport.VIOLATIONS)
    raw = nodes.view(np.uint8).reshape(-1, 60)
    written = raw.any(axis=1)
    gaps = np.nonzero(~written)[0]
    cell = nodes["cell_size"]
    fine = np.nonzero(written & (np.arange(nodes.size) > 0) & (cell >= np.float32(8.0)))[0]
    assert gaps.size and fine.size

    def plant(slot, edit, field):
        nd = nodes[slot].copy()
        edit(nd)
        _write_node(ctx, r, slot, nd)
        rep = ctx.validate_octree()
        _write_node(ctx, r, slot, nodes[slot])
        bad = nodes.copy()
        bad[slot] = nd
        with pytest.raises(AssertionError):
            check_octree(bad, *args)
        assert rep[field] == 1, (field, rep)
        return rep

    def set_cell(v):
        def f(nd):
            nd["cell_size"] = v
        return f

    def set_field(name, v):
        def f(nd):
            nd[name] = v
        return f

    def shift_corner(nd):
        nd["corner"][0] = np.float32(nd["corner"][0] + nd["cell_size"] * np.float32(0.5))

    def set_w(nd):
        nd["corner"][3] = np.float32(0.5)

    s = int(fine[0])
    rep = plant(int(gaps[0]), set_cell(np.float32(1024.0)), "written_uncovered")
    assert rep["written"] == base["written"] + 1
    rep = plant(s, lambda nd: nd.fill(0), "covered_unwritten")
    assert rep["written"] == base["written"] - 1
    plant(s, set_field("child_leaf_mask", 4), "leaf_mask")
    plant(s, shift_corner, "off_grid")
    plant(s, set_cell(np.float32(cell[s] * np.float32(3.0))), "bad_cell")
    plant(s, set_w, "bad_w")
    plant(0, set_cell(np.float32(512.0)), "bad_root")
    after = ctx.validate_octree()
    assert after == base

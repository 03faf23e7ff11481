"""Structural checks of an octree node array that hold for the reference's
semantics whatever the order of its overlapping writes, in the spirit of
checkTree (apps/cpu_demo/main.cpp:13-35).

The reference's offsets are an inclusive scan (kernels/06_partial_sum.cpp:3-15)
while ProcessOctNode (include/shared/oct.h:24-90) uses P[i] as chain i's first
slot, so chains with e >= 2 overlap the next non-empty chain and overwrite
its geometry: node-to-child geometry (a child inside its parent's octant)
does not hold in the reference's own output. What does hold, and is checked
here without restating the octree functions:

- the written slots are exactly slot 0 and the chain ranges
  [P[i], P[i] + e[i]) of the BRT nodes i >= 1 (e and P come from the edge
  count and scan, which the compiled reference pins);
- every other slot is all zero bytes (nodes start zeroed);
- every written slot has w = 1, a cell size range / 2^k for an integer k,
  and a corner on the grid of octree level root_level + k, i.e. the decode of
  a Morton prefix of that many levels (morton.h:45-80 with oct.h:56-57, 87-88);
- child_leaf_mask is 0 everywhere (SetLeaf clears bits of a zeroed word,
  oct.h:17-21);
- the root slot has the root corner and cell size = range.
"""
import numpy as np


def check_octree(nodes, codes, prefix_n, edge, offsets, min_coord=0.0, range_=1024.0):
    nodes = np.asarray(nodes)
    k = nodes.size
    u = codes.size
    assert k >= 1
    # written slots from the chain ranges
    expect = np.zeros(k, bool)
    expect[0] = True
    m = u - 1
    if m > 1:
        e = np.asarray(edge[1:m], np.int64)
        p = np.asarray(offsets[1:m], np.int64)
        live = e > 0
        starts = p[live]
        ends = p[live] + e[live]
        assert ends.max(initial=0) <= k, "a chain ends past K"
        diff = np.zeros(k + 1, np.int64)
        np.add.at(diff, starts, 1)
        np.add.at(diff, ends, -1)
        expect |= np.cumsum(diff[:k]) > 0
    raw = nodes.view(np.uint8).reshape(k, 60)
    written = raw.any(axis=1)
    bad = np.nonzero(written & ~expect)[0]
    assert bad.size == 0, f"slot {bad[:5]} written outside every chain"
    cell = nodes["cell_size"].astype(np.float64)
    corner = nodes["corner"].astype(np.float64)
    # every slot a chain covers is written (its cell size is never 0)
    missing = np.nonzero(expect & (cell == 0))[0]
    assert missing.size == 0, f"covered slot {missing[:5]} has no geometry"
    assert np.all(nodes["child_leaf_mask"] == 0), "child_leaf_mask must stay 0"
    w = expect
    assert np.all(corner[w, 3] == 1.0), "w must be 1"
    lv = np.log2(range_ / cell[w])
    assert np.all(lv == np.round(lv)) and lv.min() >= 0, "cell size not range / 2^k"
    root_level = int(prefix_n[0]) // 3 if m >= 1 else 0
    level = root_level + np.round(lv).astype(np.int64)
    assert level.max(initial=0) <= 11, "level beyond the code length"
    # The corner is the float32 decode q / 1024 * range + min of an axis value
    # q on the level's grid: some q near the real-valued inverse decodes to
    # it bit for bit (for a box other than [0, 1024) the inverse is not an
    # integer itself).
    q = np.floor((corner[w, :3] - min_coord) / range_ * 1024.0)
    step = (2 ** np.maximum(10 - level, 0))[:, None]
    f32 = np.float32
    got = nodes["corner"][w, :3]
    ok = np.zeros(got.shape, bool)
    on_any = np.zeros(got.shape, bool)
    for d in (-1, 0, 1, 2):
        cand = q + d
        inside = (cand >= 0) & (cand < 1024)
        dec = (np.clip(cand, 0, 1023).astype(f32) / f32(1024.0) * f32(range_)).astype(f32) \
            + f32(min_coord)
        hit = inside & (dec == got)
        on_any |= hit
        ok |= hit & (np.mod(cand, step) == 0)
    assert np.all(on_any), "corner off the grid"
    assert np.all(ok), "corner not on its level's grid"
    # root: corner of codes[0]'s root-level prefix, cell size = range
    sh = (30 - 3 * root_level) & 31
    c0 = (int(codes[0]) >> sh) << sh
    axes = [sum(((c0 >> (3 * b + a)) & 1) << b for b in range(10)) for a in range(3)]
    np.testing.assert_array_equal(
        nodes["corner"][0],
        np.array([np.float32(np.float32(v) / np.float32(1024.0)) * np.float32(range_)
                  + np.float32(min_coord) for v in axes] + [1.0], np.float32))
    assert nodes["cell_size"][0] == np.float32(range_)
    return int(w.sum())

"""Benchmark of the octree-construction pipeline on one MI355X.

Default workload: the plain pipeline on 10,000,000 float4 points uniform in
[0, 1024)^3 (seed 114514), 32-bit Morton codes, all seven stages on device,
replayed from a captured hipGraph with the input already resident in HBM.

One step = one whole build (Morton -> OneSweep -> unique -> radix tree ->
edge count + scan -> octree). ``value`` is N / (time per build) in Mpoints/s.

Also reported (same JSON line):
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average duration from HIP events on the context stream;
                `traffic` is the HBM bytes per launch that rocprofv3's
                FETCH_SIZE/WRITE_SIZE counters measured for that kernel on
                this workload (profiles/traffic_plain_10M.json), when present;
  e2e_roofline  the whole build's algorithmic bytes (60N + 61U + 60K) / the
                time per build, against 8 TB/s;
  kernels_ms    every kernel's average duration in eager builds;
  kernels_ms_cold  the same with 512 MiB written before every kernel (no
                input left in L2 or the Infinity Cache); kernels_gbs[_cold]
                algorithmic bytes / duration per kernel;
  float3, c1_307200  the workload as packed float3 points, and the
                307,200-point CPU-path cloud, from graphs;
  cpu_baseline  the CPU oracle (the reference CPU path restated, OpenMP) timed
                on this host in the same run, on every CPU the process may use
                and on one thread, for the workload and the 307,200-point cloud;
  stream        frames through the streaming pipeline (ocp_ppl) at depths
                1-3, from HBM and from pinned host memory;
  batched       512 clouds (log-uniform 4096..2M points, uniform or clustered,
                each with its own box) built by one segmented launch per stage
                from a hipGraph, against 512 sequential plain builds (each
                replayed from its own graph) of the same clouds;
  learned_sort  2^26 clustered points over 30 successive frames (rigid motion
                <= 0.5 deg, <= 2 units, applied on the GPU): the CDF model is
                trained on 1% key samples of the earlier frames and sorts
                frame t; OneSweep sorts the same keys; both outputs are
                compared.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain|clustered]
                  [--learned-frames F] [--no-batched] [--no-learned]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "end-to-end octree Mpoints/s + HBM roofline %, plain vs batched/learned-sort"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(name: str, n: int, u: int, k: int, float4: bool = True) -> float:
    """Algorithmic HBM bytes of one launch (per-unit figure x units)."""
    m = max(u - 1, 0)
    pt = 16 if float4 else 12
    table = {
        "morton_hist": (pt + 4) * n,            # point read + key write
        "onesweep_pass0": 8 * n,                # key read + key write
        "onesweep_pass1": 8 * n,
        "onesweep_pass2": 8 * n,
        "onesweep_pass3": 8 * n,
        "msd_scatter": 8 * n,                   # MSD sort: scatter by the top digit
        "msd_buckets": 8 * n,                   # every bucket read and written once
        "msd_lsd_fallback": 28 * n,             # lower-digit counts + three passes
        "unique": 4 * n + 4 * u,                # sorted read + unique write
        "radix_tree": 4 * u + 11 * m,           # codes read; 1+1+1+4+4 written
        "edge_scan": 14 * m,                    # 1+4+1 read; 4+4 written
        "octree": 24 * m + 60 * k,              # BRT reads + one 60-B node per slot
    }
    return float(table.get(name, 0.0))


def e2e_bytes(n: int, u: int, k: int) -> float:
    # 20N (Morton) + 36N (sort) + 4N+4U (unique) + 4U+11M (tree) + 10M (edge)
    # + 8M (scan) + 24M+60K (octree), M = U-1. The sort term is that of the
    # reference's four-pass LSD sort (a histogram read and four passes) for
    # every sort form, so the figure stays one fixed amount of work per build:
    # the MSD form moves about 16N of it.
    return 60.0 * n + 61.0 * u + 60.0 * k


def omp_threads() -> int:
    """Host threads this process may use: the CPUs it is allowed to run on,
    capped by OMP_NUM_THREADS when the environment sets it."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return min(int(v), avail)
    return avail


def cpu_info() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = None
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_available": avail,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def _cpu_run(pts, threads: int, budget_s: float):
    """Median per-stage ms of the CPU chain, repeated within a time budget."""
    import oracle

    t0 = time.perf_counter()
    ms, u, k = oracle.time_pipeline(pts, 0.0, 1024.0, nthreads=threads, reps=1)
    one = (time.perf_counter() - t0) / 2  # warm-up + one timed run
    reps = 1
    if budget_s > 3 * one:
        reps = int(max(1, min(5, (budget_s - 2 * one) / max(one, 1e-3))))
        ms, u, k = oracle.time_pipeline(pts, 0.0, 1024.0, nthreads=threads, reps=reps)
    return ms, u, k, reps


def cpu_baseline(pts: np.ndarray, c1_pts: np.ndarray, budget_s: float = 20.0):
    """The CPU oracle's full chain on this host: the workload on every
    thread the process may use and on one thread, plus the 307,200-point
    CPU-path cloud."""
    nt = omp_threads()
    n = pts.shape[0]
    ms, u, k, reps = _cpu_run(pts, nt, budget_s)
    ms1, u1, k1, reps1 = _cpu_run(pts, 1, 5.0)   # ~0.5 s a run: five runs
    c_all, _, _, c_reps = _cpu_run(c1_pts, nt, 2.0)
    c_one, _, _, c_reps1 = _cpu_run(c1_pts, 1, 2.0)
    m = c1_pts.shape[0]
    return {
        "value": n / (ms["total"] * 1e-3) / 1e6,
        "unit": "Mpoints/s",
        "cores": nt,
        "kind": "port",
        "sample": (f"full workload ({n} points), median of {reps} timed runs after "
                   f"a warm-up, OpenMP on {nt} threads (every CPU this process may use)"),
        "cpu": cpu_info(),
        "stage_ms": {kk: round(v, 3) for kk, v in ms.items()},
        "one_thread": {"value": n / (ms1["total"] * 1e-3) / 1e6, "unit": "Mpoints/s",
                       "runs": reps1, "stage_ms": {kk: round(v, 3) for kk, v in ms1.items()},
                       "counts_match": bool(u1 == u and k1 == k)},
        "c1_307200": {"points": m, "ms_all_threads": round(c_all["total"], 3),
                      "ms_one_thread": round(c_one["total"], 3),
                      "Mpts_s_all_threads": round(m / (c_all["total"] * 1e-3) / 1e6, 2),
                      "Mpts_s_one_thread": round(m / (c_one["total"] * 1e-3) / 1e6, 2)},
        "n_unique": u,
        "n_oct": k,
    }


def load_traffic(kernel: str, n: int):
    """Measured HBM bytes per launch of `kernel` for the default workload, and
    where they come from (the profile file and the commit it was taken at)."""
    path = os.path.join(ROOT, "profiles", "traffic_plain_10M.json")
    if n != 10_000_000 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        table = json.load(f)
    v = table.get("kernels", {}).get(kernel)
    src = {"file": "profiles/traffic_plain_10M.json", "commit": table.get("commit"),
           "command": table.get("command")}
    return (None if v is None else float(v)), src


def _segment(ctx, r, a: int, b: int, u: int, k: int, node0: int) -> dict:
    """Host copies of one segment's outputs (point offset a..b, node offset
    node0) from a context's results."""
    from pipeline_amd import d2h

    m = max(u - 1, 0)
    out = {"sorted": d2h(ctx, r.sorted_keys + 4 * a, b - a, np.uint32),
           "unique": d2h(ctx, r.unique_keys + 4 * a, u, np.uint32),
           "prefix_n": d2h(ctx, r.prefix_n + a, m, np.uint8),
           "has_leaf_left": d2h(ctx, r.has_leaf_left + a, m, np.uint8),
           "has_leaf_right": d2h(ctx, r.has_leaf_right + a, m, np.uint8),
           "left_child": d2h(ctx, r.left_child + 4 * a, m, np.int32),
           "parent": d2h(ctx, r.parent + 4 * a, m, np.int32),
           "edge_count": d2h(ctx, r.edge_count + 4 * a, m, np.int32),
           "offsets": d2h(ctx, r.offsets + 4 * a, m, np.int32),
           "nodes": d2h(ctx, r.nodes + 60 * node0, 60 * k, np.uint8)}
    return out


def bench_batched(ctx, pipeline_amd, clouds, steps: int, autotune: bool):
    """512 clouds in one segmented build vs 512 sequential plain builds (on a
    second context), every output of every cloud compared byte for byte."""
    pts, offs, smin, srg, kinds = clouds.batch(512, seed=114514)
    sizes = offs[1:] - offs[:-1]
    total = int(offs[-1])
    d = ctx.upload(pts)
    del pts
    tuned = ctx.autotune(d, seg_offsets=offs, seg_min=smin, seg_range=srg, reps=2) \
        if autotune else None
    ctx.set_graph_mode(True)
    ms_b = ctx.time_builds(d, 2, max(3, steps // 5), seg_offsets=offs, seg_min=smin,
                           seg_range=srg)
    u_b, k_b = ctx.counts(512)
    rb = ctx.results()
    ost = pipeline_amd.d2h(ctx, rb.oct_start, 513, np.int32)
    seq = pipeline_amd.Context(0)
    seq.set_tune(**ctx.get_tune())
    seq.set_graph_mode(True)
    # Sequential plain builds, largest first so that the workspace is sized
    # once and every cloud's graph stays valid.
    ms_seq = 0.0
    same = True
    for k in np.argsort(-sizes, kind="stable"):
        a, b = int(offs[k]), int(offs[k + 1])
        ms_seq += seq.time_builds(d, 1, 3, n=int(sizes[k]), min_coord=float(smin[k]),
                                  range_=float(srg[k]), first=a)
        u1, k1 = seq.counts(1)
        if u1[0] != u_b[k] or k1[0] != k_b[k]:
            same = False
            continue
        one = _segment(seq, seq.results(), 0, b - a, int(u1[0]), int(k1[0]), 0)
        bat = _segment(ctx, rb, a, b, int(u_b[k]), int(k_b[k]), int(ost[k]))
        same &= all(np.array_equal(one[f], bat[f]) for f in one)
    seq.close()
    d.free()
    thr_b = total / (ms_b * 1e-3) / 1e6
    thr_s = total / (ms_seq * 1e-3) / 1e6
    return {
        "clouds": 512, "points": total, "min_cloud": int(sizes.min()),
        "max_cloud": int(sizes.max()), "clustered_clouds": int(kinds.sum()),
        "ms_batched": round(ms_b, 4), "ms_sequential": round(ms_seq, 4),
        "batched_Mpts_s": round(thr_b, 2), "sequential_Mpts_s": round(thr_s, 2),
        "speedup": round(thr_b / thr_s, 3), "outputs_match_sequential": bool(same),
        "compared": "every stage output of every cloud, byte for byte",
        "tune": tuned,
    }


def bench_stream(pipeline_amd, clouds, n: int, frames: int = 32, tune: dict | None = None):
    """Frames of n points through the streaming pipeline (ocp_ppl) at depths
    1-3: frames already in HBM, and frames uploaded from pinned host memory
    (PCIe-inclusive), every slot with the plain workload's tile table. Counts
    of each slot's last frame are checked against plain builds of the same
    points."""
    srcs = [clouds.uniform(n, seed=114514 + f) for f in range(3)]
    ref = pipeline_amd.Context(0)
    expect = []
    for p in srcs:
        dp = ref.upload(p)
        ref.build(dp, n)
        u, k = ref.counts(1)
        expect.append((int(u[0]), int(k[0])))
        dp.free()
    ref.close()
    out = {"n": n, "frames": frames}
    ok = True
    for depth in (1, 2, 3):
        with pipeline_amd.Pipeline(depth, max_points=n) as ppl:
            if tune:
                for slot in range(depth):
                    ppl.context(slot).set_tune(**tune)
            dev = [ppl.context(0).upload(p) for p in srcs]
            for slot in range(depth):
                ppl.host_buffer(slot, n)[:] = srcs[slot % 3]

            def run(staged: bool, count: int) -> float:
                t0 = time.perf_counter()
                for f in range(count):
                    slot = f % depth
                    if f >= depth:
                        ppl.wait(slot)
                    ppl.submit(slot, None if staged else dev[f % 3], n)
                for slot in range(min(depth, count)):
                    ppl.wait(slot)
                return (time.perf_counter() - t0) * 1e3 / count

            def check(staged: bool, count: int) -> bool:
                good = True
                for slot in range(min(depth, count)):
                    last = max(f for f in range(count) if f % depth == slot)
                    src = slot % 3 if staged else last % 3
                    u, k = ppl.context(slot).counts(1)
                    good &= (int(u[0]), int(k[0])) == expect[src]
                return good

            run(False, 2 * depth)  # sizes the slots' workspaces
            ms_dev = run(False, frames)
            ok &= check(False, frames)
            ms_host = run(True, frames)
            ok &= check(True, frames)
            for a in dev:
                a.free()
        out[f"depth{depth}"] = {"ms_per_frame_hbm": round(ms_dev, 4),
                                "Mpts_s_hbm": round(n / (ms_dev * 1e-3) / 1e6, 1),
                                "ms_per_frame_pcie": round(ms_host, 4),
                                "Mpts_s_pcie": round(n / (ms_host * 1e-3) / 1e6, 1)}
    out["counts_match"] = bool(ok)
    out["tune"] = tune
    base = out["depth1"]
    for depth in (2, 3):
        cur = out[f"depth{depth}"]
        cur["speedup_hbm_vs_depth1"] = round(base["ms_per_frame_hbm"] / cur["ms_per_frame_hbm"], 3)
        cur["speedup_pcie_vs_depth1"] = round(base["ms_per_frame_pcie"] / cur["ms_per_frame_pcie"],
                                              3)
    return out


def _motion(seed: int, center: np.ndarray, max_deg: float = 0.5, max_shift: float = 2.0):
    """The 3x4 [R | t] of clouds.rigid_motion's step (same draws from the
    same seed): rotation by <= max_deg about a random axis through center,
    then a shift <= max_shift per axis; returns (affine, shift)."""
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.deg2rad(rng.uniform(-max_deg, max_deg))
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    rot = np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)
    shift = rng.uniform(-max_shift, max_shift, 3)
    return np.hstack([rot, (center + shift - rot @ center)[:, None]]), shift


def bench_learned(ctx, clouds, frames: int, n: int = 1 << 26, reps: int = 5):
    """Learned-bucket sort vs OneSweep over successive frames of a moving
    clustered cloud. Each frame moves on the device (ocp_transform_points,
    the rotation centre carried forward by the shifts, clipping aside) and
    its keys come from the pipeline's Morton kernel. Frame t is sorted with a
    model trained on 1% key samples of frames < t."""
    host = clouds.clustered(n, seed=114514)
    center = host[:, :3].astype(np.float64).mean(axis=0)
    pts = ctx.upload(host)
    del host
    hi = float(np.nextafter(np.float32(1024.0), np.float32(0.0)))
    work = ctx.empty(n, np.uint32)
    alt = ctx.empty(n, np.uint32)
    ref = ctx.empty(n, np.uint32)
    keys_prev = ctx.empty(n, np.uint32)
    keys_cur = ctx.empty(n, np.uint32)

    def timed(fn, src):
        ts = []
        for _ in range(reps + 1):
            ctx.copy(work, src, 4 * n)
            ctx.synchronize()
            t0 = time.perf_counter()
            fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts[1:])) * 1e3

    ctx.k_ComputeMortonCode(pts, keys_prev, n, 0.0, 1024.0)
    ctx.cdf_clear_samples()
    per = []
    for t in range(1, frames):
        affine, shift = _motion(1000 + t, center)
        center = center + shift
        ctx.transform_points(pts, pts, affine, 0.0, hi)
        ctx.k_ComputeMortonCode(pts, keys_cur, n, 0.0, 1024.0)
        ctx.cdf_add_samples(keys_prev, n, 100)     # 1% samples of frame t-1
        ctx.synchronize()
        t0 = time.perf_counter()
        loss = ctx.cdf_train(steps=2000, lr=1.0, momentum=0.9, seed=t)
        train_ms = (time.perf_counter() - t0) * 1e3
        ms_os = timed(lambda: ctx.k_SortKeys(work, alt, n), keys_cur)
        ctx.copy(ref, work, 4 * n)
        st = {}
        ms_l = timed(lambda: st.update(ctx.sort_learned(work, alt, n)), keys_cur)
        exact = bool(np.array_equal(work.to_host(), ref.to_host()))
        per.append({"frame": t, "onesweep_ms": round(ms_os, 4), "learned_ms": round(ms_l, 4),
                    "exact": exact, "train_ms": round(train_ms, 1), "loss": loss,
                    "keys_fallback": st["keys_big"], "max_bucket": st["max_bucket"],
                    "fallback_passes": st["fallback_passes"]})
        keys_prev, keys_cur = keys_cur, keys_prev
    ctx.cdf_reset(0, (1 << 30) - 1)
    ms_lin = timed(lambda: ctx.sort_learned(work, alt, n), keys_prev)
    for a in (pts, work, alt, ref, keys_prev, keys_cur):
        a.free()
    os_ms = float(np.mean([p["onesweep_ms"] for p in per]))
    l_ms = float(np.mean([p["learned_ms"] for p in per]))
    return {
        "n": n, "frames": frames, "motion": "rotation <= 0.5 deg about a random axis through "
        "the centre, shift <= 2 units per axis, per frame, applied on the GPU",
        "samples": "every 100th key of each earlier frame (the last 2^22 kept)",
        "onesweep_ms": round(os_ms, 4), "learned_ms": round(l_ms, 4),
        "speedup_vs_onesweep": round(os_ms / l_ms, 3),
        "untrained_linear_cdf_ms": round(ms_lin, 4),
        "all_exact": all(p["exact"] for p in per), "per_frame": per,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="plain", choices=["plain", "clustered"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default=None,
                    help="tile table applied without timing, e.g. "
                         "sort_bits=10,sort10_kpt=16,sort_msd=1 (the rest default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device copy (traffic-counter reference)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the per-kernel run with a cache flush before every kernel")
    ap.add_argument("--tune-cache", default=None,
                    help="tuning cache file: reuse a tuned workload class instead of timing")
    ap.add_argument("--learned-frames", type=int, default=30,
                    help="frames of the learned-sort comparison (the full 30-frame sequence)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the pipeline has no multi-GPU path in this build)", file=sys.stderr)
        return 2

    import pipeline_amd
    from pipeline_amd import clouds

    if args.workload == "plain":
        n = args.n or 10_000_000
        pts = clouds.uniform(n, seed=114514)
        wl = f"plain, {n} float4 points uniform in [0,1024)^3, 32-bit Morton"
    else:
        n = args.n or (1 << 24)
        pts = clouds.clustered(n, seed=114514)
        wl = f"plain, {n} float4 points clustered (plane/cylinders/gaussians), 32-bit Morton"

    ctx = pipeline_amd.Context(0)
    d = ctx.upload(pts)
    if args.copy_calibration:
        nb = 512 << 20
        a_buf = ctx.empty(nb // 4, np.uint32)
        b_buf = ctx.empty(nb // 4, np.uint32)
        ctx.copy(b_buf, a_buf, nb)
        ctx.synchronize()
        a_buf.free()
        b_buf.free()
    tuned = None
    if args.tune:
        ctx.set_tune(**{k: int(v) for k, v in (f.split("=") for f in args.tune.split(","))})
    elif not args.no_autotune:
        tuned = ctx.autotune(d, n=n, reps=10, cache=args.tune_cache)
    plain_tune = ctx.get_tune()
    # Eager builds with per-kernel HIP events: back to back (warm), then with
    # 512 MiB written before every kernel (cold: L2 and the 256 MB Infinity
    # Cache hold none of its input).
    def per_kernel(cold: bool) -> dict:
        if cold:
            ctx.set_cold_timing(512 << 20)
        ctx.set_stage_timing(True)
        kms: dict[str, list[float]] = {}
        for _ in range(4):
            ctx.build(d, n, 0.0, 1024.0)
            ctx.synchronize()
            for kname, v in ctx.kernel_ms().items():
                kms.setdefault(kname, []).append(v)
        ctx.set_stage_timing(False)
        if cold:
            ctx.set_cold_timing(0)
        return {kname: float(np.mean(v[1:])) for kname, v in kms.items()}

    print("per-kernel timing: warm, cold", file=sys.stderr, flush=True)
    kmean = per_kernel(False)
    kcold = {} if args.no_cold else per_kernel(True)
    u_arr, k_arr = ctx.counts(1)
    u, k = int(u_arr[0]), int(k_arr[0])

    def gbs(table: dict) -> dict:
        return {kk: round(kernel_bytes(kk, n, u, k) / (v * 1e-3) / 1e9, 1)
                for kk, v in table.items() if kernel_bytes(kk, n, u, k) > 0 and v > 0}

    print(f"timed: {args.steps} graph replays", file=sys.stderr, flush=True)
    # Timed region: K graph replays, synchronised on both sides.
    ctx.set_graph_mode(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ms_per = ctx.time_builds(d, args.warmup, args.steps, n=n)
    wall = time.perf_counter() - t0
    ctx.synchronize()
    u_arr, k_arr = ctx.counts(1)
    assert int(u_arr[0]) == u and int(k_arr[0]) == k

    # The same cloud as packed float3 (12 B per point), and the 307,200-point
    # CPU-path cloud, both from graphs.
    d3 = ctx.upload(np.ascontiguousarray(pts[:, :3]))
    ms_f3 = ctx.time_builds(d3, args.warmup, args.steps, n=n)
    d3.free()
    c1 = clouds.uniform(307_200, seed=114514)
    dc1 = ctx.upload(c1)
    ms_c1 = ctx.time_builds(dc1, args.warmup, args.steps, n=c1.shape[0])
    dc1.free()

    value = n / (ms_per * 1e-3) / 1e6
    dom = max((kk for kk in kmean if kernel_bytes(kk, n, u, k) > 0), key=lambda kk: kmean[kk])
    dom_bytes = kernel_bytes(dom, n, u, k)
    achieved = dom_bytes / (kmean[dom] * 1e-3) / 1e9
    eb = e2e_bytes(n, u, k)
    e2e = eb / (ms_per * 1e-3) / 1e9
    traffic, traffic_src = load_traffic(dom, n)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": wl, "n_points": n, "n_unique": u, "n_oct": k,
                   "seed": 114514, "hipgraph": True, "tune": ctx.get_tune(),
                   "autotuned": tuned is not None,
                   "tune_from_cache": bool(getattr(ctx, "tune_from_cache", False))},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_launch": dom_bytes, "avg_ms": round(kmean[dom], 5),
                     "cold_avg_ms": round(kcold.get(dom, 0.0), 5),
                     "cold_frac": round(dom_bytes / (kcold[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                     if kcold.get(dom) else None},
        "e2e_roofline": {"achieved": round(e2e, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(e2e / HBM_PEAK_GBS, 4), "bytes": eb},
        "kernels_ms": {kk: round(v, 5) for kk, v in kmean.items()},
        "kernels_gbs": gbs(kmean),
        "kernels_ms_cold": {kk: round(v, 5) for kk, v in kcold.items()},
        "kernels_gbs_cold": gbs(kcold),
        "float3": {"ms_per_build": round(ms_f3, 4),
                   "Mpts_s": round(n / (ms_f3 * 1e-3) / 1e6, 2)},
        "c1_307200": {"ms_per_build": round(ms_c1, 4),
                      "Mpts_s": round(c1.shape[0] / (ms_c1 * 1e-3) / 1e6, 2)},
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_cpu:
        print("cpu baseline: all threads, one thread, 307,200 points", file=sys.stderr,
              flush=True)
        cb = cpu_baseline(pts, c1)
        cb_ok = cb["n_unique"] == u and cb["n_oct"] == k
        line["cpu_baseline"] = cb
        line["cpu_counts_match"] = bool(cb_ok)
    d.free()
    del pts
    if not args.no_batched:
        print("batched: 512 clouds", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(ctx, pipeline_amd, clouds, args.steps,
                                        not args.no_autotune)
    if not args.no_stream:
        print("streaming: 3 depths", file=sys.stderr, flush=True)
        line["stream"] = bench_stream(pipeline_amd, clouds, n, tune=plain_tune)
    if not args.no_learned and args.learned_frames >= 2:
        print(f"learned sort: {args.learned_frames} frames", file=sys.stderr, flush=True)
        line["learned_sort"] = bench_learned(ctx, clouds, args.learned_frames)
    ctx.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Learned-CDF bucket sort: the model's gradient against the float64 oracle
(itself pinned by finite differences), the compiled bucket boundaries against
the oracle's, the two bucket evaluations against each other, and the sort
against np.sort on every input shape, including models that bucket badly."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import cdf
from pipeline_amd import clouds

KMAX30 = (1 << 30) - 1


def morton_keys(pts: np.ndarray) -> np.ndarray:
    return oracle.morton(pts, 0.0, 1024.0, nthreads=8)


def random_root(rng) -> np.ndarray:
    p = cdf.linear_params(0, KMAX30)
    p[cdf.W1:cdf.W1 + 16] = rng.uniform(0.2, 2.0, 16)
    p[cdf.B1:cdf.B1 + 16] = rng.uniform(-1.0, 0.2, 16)
    p[cdf.W2:cdf.W2 + 16] = rng.uniform(0.0, 1.0, 16)
    p[cdf.B2] = rng.uniform(-0.1, 0.1)
    return p


# ---------------------------------------------------------------- oracle (CPU)
def test_oracle_grad_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_root(rng).astype(np.float64)[:cdf.ROOT_PARAMS]
    x = rng.random(512)
    t = rng.random(512)
    g = cdf.root_grad(p, x, t)
    fd = cdf.finite_diff_grad(p, x, t, eps=1e-6)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)


def test_oracle_identity_root_and_linear_buckets():
    p = cdf.linear_params(0, KMAX30)
    x = np.linspace(0.0, 0.999, 101)
    np.testing.assert_allclose(cdf.root_forward(p, x), x, atol=1e-7)
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 1 << 30, 4096, dtype=np.uint64).astype(np.uint32)
    kleaf = cdf.leaf_starts(p, 0, KMAX30)
    q = cdf.q_model(p, 0, KMAX30, kleaf, keys)
    ideal = (keys.astype(np.float64) * cdf.BUCKETS / (1 << 30)).astype(np.int64)
    assert np.max(np.abs(q - ideal)) <= 1


def test_oracle_boundaries_reproduce_the_model():
    rng = np.random.default_rng(5)
    p = random_root(rng)
    kmin, kmax = 1000, KMAX30
    fb = cdf.boundaries(p, kmin, kmax)
    assert np.all(np.diff(fb.astype(np.int64)) >= 0)
    kleaf = cdf.leaf_starts(p, kmin, kmax)
    probe = np.concatenate([rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32),
                            fb, fb - 1, np.array([0, 0xFFFFFFFF], np.uint32)])
    probe = np.sort(probe)
    q = cdf.q_model(p, kmin, kmax, kleaf, probe)
    assert np.all(np.diff(q) >= 0)
    np.testing.assert_array_equal(cdf.bucket_of(fb, probe), q)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(11)
    p = random_root(rng)
    ctx.cdf_set(p, 0, KMAX30)
    x = rng.random(8192).astype(np.float32)
    t = rng.random(8192).astype(np.float32)
    g = ctx.cdf_grad(x, t)
    go = cdf.root_grad(p[:cdf.ROOT_PARAMS], x, t)
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-6)


@pytest.mark.gpu
def test_gpu_linear_model_matches_oracle(ctx):
    kmin, kmax = 12345, (1 << 30) - 77
    ctx.cdf_reset(kmin, kmax)
    p, lo, hi = ctx.cdf_get()
    assert (lo, hi) == (kmin, kmax)
    np.testing.assert_array_equal(p, cdf.linear_params(kmin, kmax))
    np.testing.assert_array_equal(ctx.cdf_boundaries(), cdf.boundaries(p, kmin, kmax))


def _train(ctx, keys_host, stride=8, steps=1500, seed=1):
    d = ctx.upload(keys_host)
    ctx.cdf_clear_samples()
    ctx.cdf_add_samples(d, keys_host.size, stride)
    loss = ctx.cdf_train(steps=steps, lr=1.0, momentum=0.9, seed=seed)
    d.free()
    return loss


@pytest.mark.gpu
def test_gpu_training_and_compiled_boundaries(ctx):
    keys = morton_keys(clouds.clustered(1 << 20, seed=21))
    loss = _train(ctx, keys)
    assert np.isfinite(loss) and loss < 2e-3  # the identity start is ~3e-2
    p, kmin, kmax = ctx.cdf_get()
    s = np.sort(keys[::8])  # the key range of the samples
    assert (kmin, kmax) == (int(s[0]), int(s[-1]))
    w1, w2 = p[cdf.W1:cdf.W1 + 16], p[cdf.W2:cdf.W2 + 16]
    assert np.all(w1 * w2 >= 0)  # every hidden unit non-decreasing
    lo = p[cdf.LEAF_LO:cdf.LEAF_LO + cdf.LEAVES + 1]
    assert lo[-1] == 1.0 and np.all(np.diff(lo) >= 0)
    assert np.all(p[cdf.LEAF_B:cdf.LEAF_B + cdf.LEAVES] >= 0)
    fb = ctx.cdf_boundaries()
    np.testing.assert_array_equal(fb, cdf.boundaries(p, kmin, kmax))
    # model evaluation, compiled evaluation and the oracle agree key by key
    probe = np.concatenate([keys[:200000], fb, fb - 1,
                            np.array([0, kmin, kmax, 0xFFFFFFFF], np.uint32)])
    d = ctx.upload(probe)
    q0 = ctx.cdf_buckets(d, mode=0)
    q1 = ctx.cdf_buckets(d, mode=1)
    q2 = ctx.cdf_buckets(d, mode=2)  # through the cell table, as the sort does
    d.free()
    np.testing.assert_array_equal(q0, q1)
    np.testing.assert_array_equal(q0, q2)
    kleaf = cdf.leaf_starts(p, kmin, kmax)
    np.testing.assert_array_equal(q0, cdf.q_model(p, kmin, kmax, kleaf, probe))


@pytest.mark.gpu
def test_gpu_training_is_deterministic(ctx):
    keys = morton_keys(clouds.clustered(1 << 18, seed=22))
    _train(ctx, keys, steps=300, seed=7)
    a = ctx.cdf_get()[0]
    _train(ctx, keys, steps=300, seed=7)
    b = ctx.cdf_get()[0]
    np.testing.assert_array_equal(a, b)


def _sort_cases():
    rng = np.random.default_rng(31)
    return {
        "clustered_2m": morton_keys(clouds.clustered(1 << 21, seed=32)),
        "uniform_1m": morton_keys(clouds.uniform(1 << 20, seed=33)),
        "full_range_u32": rng.integers(0, 1 << 32, 300000, dtype=np.uint64).astype(np.uint32),
        "few_values": rng.choice(np.array([5, 77, 1 << 29, 0xFFFFFFFF], np.uint32), 200000),
        "all_equal": np.full(100000, 12345, np.uint32),
        "sorted": np.arange(150000, dtype=np.uint32) * 7,
        "reversed": (np.arange(150000, dtype=np.uint32) * 7)[::-1].copy(),
        "two": np.array([9, 3], np.uint32),
        "odd_small": rng.integers(0, 1 << 30, 1001, dtype=np.uint64).astype(np.uint32),
    }


@pytest.mark.gpu
def test_gpu_sort_learned_bit_exact(ctx):
    train = morton_keys(clouds.clustered(1 << 20, seed=30))
    _train(ctx, train)
    for name, h in _sort_cases().items():
        print(f"case {name} n={h.size}", flush=True)
        d = ctx.upload(h)
        alt = ctx.empty(h.size, np.uint32)
        st = ctx.sort_learned(d, alt, h.size)
        np.testing.assert_array_equal(d.to_host(), np.sort(h), err_msg=name)
        assert st["misordered"] == 0 and st["full_resort"] == 0, name
        d.free()
        alt.free()


@pytest.mark.gpu
def test_gpu_sort_learned_bad_models_fall_back(ctx):
    h = morton_keys(clouds.clustered(1 << 20, seed=40))
    for name, p in (("zeros", np.zeros(cdf.PARAMS, np.float32)),
                    ("nan", np.full(cdf.PARAMS, np.nan, np.float32))):
        print(f"model {name}", flush=True)
        ctx.cdf_set(p, 0, KMAX30)
        d = ctx.upload(h)
        alt = ctx.empty(h.size, np.uint32)
        st = ctx.sort_learned(d, alt, h.size)
        np.testing.assert_array_equal(d.to_host(), np.sort(h), err_msg=name)
        # every key lands in one bucket, which the OneSweep fallback sorts
        assert st["n_big"] == 1 and st["keys_big"] == h.size, (name, st)
        d.free()
        alt.free()


@pytest.mark.gpu
def test_gpu_fallback_sorts_only_varying_digits(ctx):
    """The OneSweep fallback sorts each oversized bucket only over the digits
    its key range [fb[q], fb[q+1]) can vary in: two passes for the buckets
    whose range varies in the low two digits only, four for the others, each
    group in place. The count is recomputed here from the model's own
    boundaries, and the output still equals np.sort."""
    ctx.cdf_reset(0, KMAX30)
    fb = ctx.cdf_boundaries().astype(np.int64)
    rng = np.random.default_rng(7)
    aligned = (int(fb[1]) + 100, 4000)         # bucket 1 is exactly [2^16, 2^17)
    straddling = (int(fb[10000]) + 10, 60000)  # bucket 10000 spans a 2^20 boundary
    for name, ranges, want in (("aligned bucket", [aligned], 2),
                               ("straddling bucket", [straddling], 4),
                               ("one of each", [aligned, straddling], 6)):
        h = np.concatenate([lo + rng.integers(0, span, 50_000) for lo, span in ranges])
        h = rng.permutation(h).astype(np.uint32)
        counts = np.bincount(cdf.bucket_of(fb.astype(np.uint32), h), minlength=fb.size)
        groups = set()
        for q in np.nonzero(counts > 16384)[0]:
            q_lo = 0 if q == 0 else int(fb[q])
            q_hi = int(fb[q + 1]) if q + 1 < fb.size else 0
            span = (q_lo ^ (q_hi - 1)) if (q + 1 < fb.size and q_hi > q_lo) else 0xFFFFFFFF
            groups.add(4 if span >> 16 else 2)
        assert sum(groups) == want, (name, groups)
        d = ctx.upload(h)
        alt = ctx.empty(h.size, np.uint32)
        st = ctx.sort_learned(d, alt, h.size)
        np.testing.assert_array_equal(d.to_host(), np.sort(h), err_msg=name)
        assert st["n_big"] == len(ranges) and st["fallback_passes"] == want, (name, st)
        d.free()
        alt.free()


@pytest.mark.gpu
def test_gpu_trained_model_beats_linear_cdf(ctx):
    """What the trained network gains over the untrained (linear-CDF)
    bucketing on clustered keys: fewer keys left to the fallback sort and a
    smaller largest bucket, both sorts exact."""
    frame0 = clouds.clustered(1 << 22, seed=50)
    frame1 = clouds.rigid_motion(frame0, seed=51)
    k0, k1 = morton_keys(frame0), morton_keys(frame1)
    ref = np.sort(k1)
    d = ctx.upload(k1)
    alt = ctx.empty(k1.size, np.uint32)
    ctx.cdf_reset(int(k0.min()), int(k0.max()))
    st_lin = ctx.sort_learned(d, alt, k1.size)
    np.testing.assert_array_equal(d.to_host(), ref)
    _train(ctx, k0, stride=100, steps=2000)
    d.copy_from(k1)
    st_tr = ctx.sort_learned(d, alt, k1.size)
    np.testing.assert_array_equal(d.to_host(), ref)
    print("linear", st_lin, "trained", st_tr, flush=True)
    assert st_tr["keys_big"] < st_lin["keys_big"]
    assert st_tr["max_bucket"] < st_lin["max_bucket"]


@pytest.mark.gpu
def test_gpu_cdf_save_load_roundtrip(ctx, tmp_path):
    keys = morton_keys(clouds.clustered(1 << 18, seed=60))
    _train(ctx, keys, steps=200)
    p, kmin, kmax = ctx.cdf_get()
    fb = ctx.cdf_boundaries()
    path = str(tmp_path / "model.cdf")
    ctx.cdf_save(path)
    raw = open(path, "rb").read()
    assert raw[:8] == b"OCPCDF1\0" and len(raw) == 32 + 4 * cdf.PARAMS
    hdr = np.frombuffer(raw[8:32], "<u4")
    assert list(hdr) == [1, cdf.PARAMS, kmin, kmax, 16, 4096]
    np.testing.assert_array_equal(np.frombuffer(raw[32:], "<f4"), p)
    ctx.cdf_reset(0, KMAX30)
    ctx.cdf_load(path)
    p2, kmin2, kmax2 = ctx.cdf_get()
    np.testing.assert_array_equal(p2, p)
    assert (kmin2, kmax2) == (kmin, kmax)
    np.testing.assert_array_equal(ctx.cdf_boundaries(), fb)
    bad = tmp_path / "bad.cdf"
    bad.write_bytes(raw[:100])
    with pytest.raises(Exception):
        ctx.cdf_load(str(bad))


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_gpu_sort_learned_2p26_moving_frames(ctx):
    """The learned-sort workload at its real size: 2^26 clustered points,
    three successive frames of small rigid motions. Frame t's keys are
    sorted with the model trained on 1% samples of the frames before it and
    compared with numpy's sort and with OneSweep's bytes; the trained model
    leaves fewer keys to the fallback than the linear CDF."""
    n = 1 << 26
    cur = clouds.clustered(n, seed=114514)
    dk = ctx.empty(n, np.uint32)
    alt = ctx.empty(n, np.uint32)
    ref = ctx.empty(n, np.uint32)
    ctx.cdf_clear_samples()
    prev = None
    try:
        for t in range(3):
            if t:
                cur = clouds.rigid_motion(cur, seed=1000 + t)
            keys = morton_keys(cur)
            if prev is not None:
                ctx.cdf_add_samples(prev, n, 100)
                _ = ctx.cdf_train(steps=2000, lr=1.0, momentum=0.9, seed=t)
                dk.copy_from(keys)
                st = ctx.sort_learned(dk, alt, n)
                got = dk.to_host()
                np.testing.assert_array_equal(got, np.sort(keys), err_msg=f"frame {t}")
                ref.copy_from(keys)
                ctx.k_SortKeys(ref, alt, n)
                ctx.synchronize()
                assert np.array_equal(ref.to_host(), got), f"frame {t} vs OneSweep"
                assert st["misordered"] == 0 and st["full_resort"] == 0, st
                if t == 2:
                    ctx.cdf_reset(int(keys.min()), int(keys.max()))
                    dk.copy_from(keys)
                    st_lin = ctx.sort_learned(dk, alt, n)
                    assert np.array_equal(dk.to_host(), got)
                    print("trained", st, "linear", st_lin, flush=True)
                    assert st["keys_big"] < st_lin["keys_big"]
            if prev is not None:
                prev.free()
            prev = ctx.upload(keys)
    finally:
        for a in (dk, alt, ref, prev):
            if a is not None:
                a.free()

"""CPU restatement of the learned-CDF bucket model (test infrastructure only).

The reference has no learned sort (it sorts with OneSweep,
apps/gpu_demo/main.cu:63-125); this module restates the model that the GPU
library defines, so that its kernels can be checked:

- ``root_forward`` / ``root_grad``: the 16-unit ReLU root and the gradient of
  the mean of 0.5 * (r(x) - t)^2, in float64. The gradient is pinned by
  central finite differences of ``loss`` (tests/test_learned.py).
- ``leaf_starts``, ``q_model``, ``boundaries``: the bucket function in float32
  with one rounding per operation (the library is built without contraction),
  and its compiled form, the 16384 bucket boundaries
  fb[f] = min{key : q(key) >= f}, found by the same binary search over the
  32-bit key space.

Parameter layout (float32): w1[16] b1[16] w2[16] b2 | lo[4097] | a[4096] | b[4096].
"""
from __future__ import annotations

import numpy as np

HIDDEN = 16
W1, B1, W2, B2 = 0, 16, 32, 48
ROOT_PARAMS = 49
LEAVES = 4096
LEAF_LO = 49
LEAF_A = LEAF_LO + LEAVES + 1
LEAF_B = LEAF_A + LEAVES
PARAMS = LEAF_B + LEAVES
BUCKETS = 16384
KEY_MAX = 0xFFFFFFFF


# ---------------------------------------------------------------- float64 root
def root_forward(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    p = np.asarray(p, np.float64)
    x = np.asarray(x, np.float64)
    z = np.outer(x, p[W1:W1 + HIDDEN]) + p[B1:B1 + HIDDEN]
    return np.maximum(z, 0.0) @ p[W2:W2 + HIDDEN] + p[B2]


def loss(p: np.ndarray, x: np.ndarray, t: np.ndarray) -> float:
    r = root_forward(p, x)
    return float(np.mean(0.5 * (r - np.asarray(t, np.float64)) ** 2))


def root_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray) -> np.ndarray:
    """d loss / d (w1, b1, w2, b2), float64."""
    p = np.asarray(p, np.float64)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.size
    z = np.outer(x, p[W1:W1 + HIDDEN]) + p[B1:B1 + HIDDEN]
    h = np.maximum(z, 0.0)
    err = h @ p[W2:W2 + HIDDEN] + p[B2] - t
    active = (z > 0.0).astype(np.float64)
    dh = err[:, None] * p[W2:W2 + HIDDEN][None, :] * active
    g = np.empty(ROOT_PARAMS, np.float64)
    g[W1:W1 + HIDDEN] = (dh * x[:, None]).sum(0) / n
    g[B1:B1 + HIDDEN] = dh.sum(0) / n
    g[W2:W2 + HIDDEN] = (err[:, None] * h).sum(0) / n
    g[B2] = err.sum() / n
    return g


def finite_diff_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, eps: float = 1e-6) -> np.ndarray:
    p = np.asarray(p, np.float64).copy()
    g = np.empty(ROOT_PARAMS, np.float64)
    for i in range(ROOT_PARAMS):
        old = p[i]
        p[i] = old + eps
        lp = loss(p, x, t)
        p[i] = old - eps
        lm = loss(p, x, t)
        p[i] = old
        g[i] = (lp - lm) / (2 * eps)
    return g


# ---------------------------------------------------------------- float32 bucket function
def _f32(v):
    return np.float32(v)


def _inv_span(kmin: int, kmax: int) -> np.float32:
    return np.float32(1.0) / (np.float32(np.uint32(kmax - kmin)) + np.float32(1.0))


def _root32(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    r = np.full(x.shape, p[B2], np.float32)
    for j in range(HIDDEN):
        z = p[W1 + j] * x + p[B1 + j]          # two roundings, as built
        h = np.fmax(z, np.float32(0.0))
        r = r + p[W2 + j] * h
    return r


def leaf_of(p: np.ndarray, kmin: int, kmax: int, keys: np.ndarray) -> np.ndarray:
    p = np.asarray(p, np.float32)
    keys = np.asarray(keys, np.uint32)
    d = np.where(keys > np.uint32(kmin), keys - np.uint32(kmin), np.uint32(0)).astype(np.uint32)
    x = d.astype(np.float32) * _inv_span(kmin, kmax)
    s = np.floor(_root32(p, x) * np.float32(LEAVES))
    out = np.zeros(keys.shape, np.int64)
    ok = s > 0
    out[ok] = np.minimum(s[ok], LEAVES - 1).astype(np.int64)
    out[keys == np.uint32(KEY_MAX)] = LEAVES - 1
    return out


def _search_min_key(pred, count: int) -> np.ndarray:
    """For i in [0, count): min key in [0, 2^32-1] with pred(keys, i) true,
    pred monotone in the key and true at 2^32-1; 32 vectorised halvings."""
    lo = np.zeros(count, np.uint64)
    hi = np.full(count, KEY_MAX, np.uint64)
    idx = np.arange(count)
    while True:
        act = lo < hi
        if not act.any():
            return lo.astype(np.uint32)
        mid = lo + ((hi - lo) >> np.uint64(1))
        ok = pred(mid.astype(np.uint32), idx)
        hi = np.where(act & ok, mid, hi)
        lo = np.where(act & ~ok, mid + np.uint64(1), lo)


def leaf_starts(p: np.ndarray, kmin: int, kmax: int) -> np.ndarray:
    """K[l] = min{key : leaf(key) >= l}, K[0] = 0, K[4096] = 2^32 - 1."""
    k = np.empty(LEAVES + 1, np.uint32)
    k[0] = 0
    k[LEAVES] = KEY_MAX
    k[1:LEAVES] = _search_min_key(
        lambda keys, i: leaf_of(p, kmin, kmax, keys) >= i + 1, LEAVES - 1)
    return k


def q_model(p: np.ndarray, kmin: int, kmax: int, kleaf: np.ndarray, keys: np.ndarray) -> np.ndarray:
    p = np.asarray(p, np.float32)
    keys = np.asarray(keys, np.uint32)
    l = leaf_of(p, kmin, kmax, keys)
    d = (keys - kleaf[l]).astype(np.uint32).astype(np.float32)
    y = p[LEAF_A + l] + p[LEAF_B + l] * d
    y = np.fmin(np.fmax(y, p[LEAF_LO + l]), p[LEAF_LO + l + 1])
    s = np.floor(y * np.float32(BUCKETS))
    out = np.zeros(keys.shape, np.int64)
    ok = s > 0
    out[ok] = np.minimum(s[ok], BUCKETS - 1).astype(np.int64)
    out[keys == np.uint32(KEY_MAX)] = BUCKETS - 1
    return out


def boundaries(p: np.ndarray, kmin: int, kmax: int) -> np.ndarray:
    """fb[f] = min{key : q(key) >= f}, then a running maximum."""
    kleaf = leaf_starts(p, kmin, kmax)
    fb = np.empty(BUCKETS, np.uint32)
    fb[0] = 0
    fb[1:] = _search_min_key(
        lambda keys, i: q_model(p, kmin, kmax, kleaf, keys) >= i + 1, BUCKETS - 1)
    return np.maximum.accumulate(fb)


def bucket_of(fb: np.ndarray, keys: np.ndarray) -> np.ndarray:
    """#{f >= 1 : fb[f] <= key}."""
    return np.searchsorted(fb[1:], np.asarray(keys, np.uint32), side="right")


def linear_params(kmin: int, kmax: int) -> np.ndarray:
    """The untrained (linear-CDF) model: identity root, uniform leaves."""
    p = np.zeros(PARAMS, np.float32)
    j = np.arange(HIDDEN, dtype=np.float32)
    up = j < HIDDEN // 2
    p[W1:W1 + HIDDEN] = np.where(up, 1.0, -1.0)
    p[B1:B1 + HIDDEN] = np.where(up, -j / np.float32(8.0), (j - np.float32(7.5)) / np.float32(8.0))
    p[W2] = 1.0
    kleaf = leaf_starts(p, kmin, kmax).astype(np.uint64)
    ell = np.arange(LEAVES + 1)
    p[LEAF_LO:LEAF_LO + LEAVES + 1] = (ell.astype(np.float32) / np.float32(LEAVES))
    k1 = np.append(kleaf[1:LEAVES], np.uint64(kmax) + np.uint64(1))
    span = np.where(k1 > kleaf[:LEAVES], (k1 - kleaf[:LEAVES]).astype(np.float64), 1.0)
    p[LEAF_A:LEAF_A + LEAVES] = ell[:LEAVES].astype(np.float32) / np.float32(LEAVES)
    p[LEAF_B:LEAF_B + LEAVES] = (1.0 / (LEAVES * span)).astype(np.float32)
    return p

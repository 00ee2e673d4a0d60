"""Shared input generators and parity assertions for the tests."""
import numpy as np


def rand_points(n, D, seed, dup_frac=0.0):
    rng = np.random.default_rng(seed)
    p = rng.random((n, D), dtype=np.float32)
    if dup_frac and n > 1:
        k = int(n * dup_frac)
        src = rng.integers(0, n, size=k)
        dst = rng.integers(0, n, size=k)
        p[dst] = p[src]
    return np.ascontiguousarray(p)


def rand_boxes(n, D, seed, mean_side=0.01):
    rng = np.random.default_rng(seed)
    lo = rng.random((n, D), dtype=np.float32)
    side = rng.exponential(mean_side, size=(n, D)).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([lo, lo + side], axis=1).astype(np.float32))


def rand_queries(q, D, seed, objs=None):
    """A mix of query shapes: random boxes of log-uniform side, boxes centred on
    data points, degenerate point boxes exactly on data points (closed-interval
    edges), inverted boxes (no hits), and the whole space."""
    rng = np.random.default_rng(seed)
    side = (10.0 ** rng.uniform(-3, -0.3, size=(q, 1))).astype(np.float32)
    c = rng.random((q, D), dtype=np.float32)
    if objs is not None and len(objs):
        pick = rng.integers(0, len(objs), size=q)
        on = rng.random(q) < 0.3
        c[on] = objs[pick[on], :D]
    h = np.float32(0.5) * side
    lo, hi = c - h, c + h
    if objs is not None and len(objs) and q >= 8:
        # exact point boxes on data points
        k = max(1, q // 16)
        pick = rng.integers(0, len(objs), size=k)
        lo[:k] = objs[pick, :D]
        hi[:k] = objs[pick, :D]
    if q >= 4:
        lo[-1], hi[-1] = np.float32(-1.0), np.float32(2.0)     # everything
        lo[-2], hi[-2] = np.float32(0.6), np.float32(0.4)      # inverted: nothing
    return np.ascontiguousarray(np.concatenate([lo, hi], axis=1).astype(np.float32))


def assert_bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, "%s: shape %s vs %s" % (what, a.shape, b.shape)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    bad = np.nonzero(a != b)
    assert bad[0].size == 0, "%s: %d mismatches, first at %s" % (
        what, bad[0].size, tuple(int(x[0]) for x in bad))


def assert_results_equal(got, exp, what, ids=True):
    assert_bits_equal(got["counts"], exp["counts"], what + " counts")
    if "hashes" in got and "hashes" in exp:
        assert_bits_equal(got["hashes"], exp["hashes"], what + " hashes")
    if ids:
        assert_bits_equal(got["offsets"], exp["offsets"], what + " offsets")
        assert_bits_equal(got["ids"], exp["ids"], what + " ids")

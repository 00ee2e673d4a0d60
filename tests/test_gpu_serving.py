"""Serving (continuous batching): each query's result is the same alone, in a
served mix of overlapping batches, and in the CPU oracle; admission by
result-buffer budget defers batches without changing results."""
import numpy as np
import pytest

from helpers import assert_bits_equal, rand_points

pytestmark = pytest.mark.gpu


def _queries(qcp, n, D, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, D), dtype=np.float32)
    s = (10.0 ** rng.uniform(-7, -3, size=n)) ** (1.0 / D)
    return qcp.boxes_from(c, np.repeat(s.astype(np.float32)[:, None], D, 1))


@pytest.fixture(scope="module")
def served(gpu, O):
    pts = rand_points(400000, 3, seed=42)
    ix = gpu.Index.build(pts, 3, fanout=64)
    t = O.build_packed(pts, 3, False, 64, O.KEY_REF)
    q = _queries(gpu, 3000, 3, 7)
    exp = t.search_dfs(q)
    return ix, q, exp


def test_served_mix_equals_alone_and_oracle(gpu, served):
    ix, q, exp = served
    arr = gpu.poisson_arrivals(len(q), 20000.0, seed=5)
    with gpu.Server(ix, slots=4, max_batch=256) as sv:
        counts, hashes, lat = sv.replay(q, arr)
        st = sv.stats()
    assert_bits_equal(counts, exp["counts"], "served counts")
    assert_bits_equal(hashes, exp["hashes"], "served hashes")
    assert st["completed"] == len(q) and st["batches"] >= 2
    assert (lat > 0).all()
    # each query alone, one at a time through the same server path
    with gpu.Server(ix, slots=1, max_batch=1) as sv1:
        for i in range(0, len(q), 97):
            c, h, _ = sv1.wait(sv1.submit(q[i]))
            assert c == exp["counts"][i] and h == exp["hashes"][i]
    # and through the batch API
    alone = ix.search(q[:50])
    assert_bits_equal(alone["counts"], exp["counts"][:50], "batch api counts")


def test_burst_overlaps_batches(gpu, served):
    ix, q, exp = served
    arr = np.zeros(len(q))           # everything arrives at once
    with gpu.Server(ix, slots=4, max_batch=128) as sv:
        counts, hashes, _ = sv.replay(q, arr)
        st = sv.stats()
    assert st["max_inflight_seen"] >= 2, st
    assert_bits_equal(counts, exp["counts"], "burst counts")
    assert_bits_equal(hashes, exp["hashes"], "burst hashes")


def test_budget_defers_admission(gpu, served):
    ix, q, exp = served
    arr = np.zeros(len(q))
    # 12 bytes of results per query: a 1 KiB budget admits ~85 queries at a time
    with gpu.Server(ix, slots=4, max_batch=80, result_budget=1024) as sv:
        counts, hashes, _ = sv.replay(q, arr)
        st = sv.stats()
    assert st["deferred_by_budget"] > 0, st
    assert_bits_equal(counts, exp["counts"], "budget counts")
    assert_bits_equal(hashes, exp["hashes"], "budget hashes")


def test_served_with_containment(gpu, served):
    """Subtrees inside a query reported from their rank ranges: same results."""
    ix, q, exp = served
    arr = gpu.poisson_arrivals(len(q), 50000.0, seed=6)
    with gpu.Server(ix, slots=4, max_batch=256, contain=True) as sv:
        counts, hashes, _ = sv.replay(q, arr)
    assert_bits_equal(counts, exp["counts"], "served counts, contain")
    assert_bits_equal(hashes, exp["hashes"], "served hashes, contain")

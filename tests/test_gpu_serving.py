"""Serving (continuous batching): each query's result (count + hash, or the
sorted ID list) is the same alone, in a served mix of overlapping batches, and
in the CPU oracle; admission by result-buffer budget defers batches without
changing results; the result table recycles its entries."""
import numpy as np
import pytest

from helpers import assert_bits_equal, rand_points

pytestmark = pytest.mark.gpu


def _queries(qcp, n, D, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, D), dtype=np.float32)
    s = (10.0 ** rng.uniform(-7, -3, size=n)) ** (1.0 / D)
    return qcp.boxes_from(c, np.repeat(s.astype(np.float32)[:, None], D, 1))


_TREE = {}


def served_tree(O):
    """The oracle tree of the served index (built once)."""
    if "t" not in _TREE:
        _TREE["t"] = O.build_packed(rand_points(400000, 3, seed=42), 3, False, 64, O.KEY_REF)
    return _TREE["t"]


@pytest.fixture(scope="module")
def served(gpu, O):
    pts = rand_points(400000, 3, seed=42)
    ix = gpu.Index.build(pts, 3, fanout=64)
    t = served_tree(O)
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


def test_served_id_lists_mix_alone_oracle(gpu, O, served):
    """ID-list service: each request's sorted rank-ID list is the same in a
    served mix, alone, and in the oracle."""
    ix, q, exp = served
    t = served_tree(O)
    ref = t.search_dfs(q, want_ids=True)
    arr = gpu.poisson_arrivals(len(q), 20000.0, seed=8)
    with gpu.Server(ix, slots=4, max_batch=256, ids=True) as sv:
        c, h, lat, off, ids = sv.replay(q, arr, ids_cap=int(ref["offsets"][-1]))
        st = sv.stats()
    assert_bits_equal(c, ref["counts"], "served ID counts")
    assert_bits_equal(h, ref["hashes"], "served ID hashes")
    assert_bits_equal(off, ref["offsets"], "served ID offsets")
    assert_bits_equal(ids, ref["ids"], "served ID lists")
    assert st["ids_served"] == int(ref["offsets"][-1]) and st["completed"] == len(q)
    with gpu.Server(ix, slots=1, max_batch=1, ids=True) as sv1:
        for i in range(0, len(q), 101):
            tk = sv1.submit(q[i])
            cnt, hsh, _ = sv1.wait(tk)
            got = sv1.result_ids(tk)
            e0, e1 = int(ref["offsets"][i]), int(ref["offsets"][i + 1])
            assert cnt == e1 - e0 and hsh == ref["hashes"][i]
            assert_bits_equal(got, ref["ids"][e0:e1], "alone IDs %d" % i)
            with pytest.raises(gpu.QcpError):
                sv1.result_ids(tk)  # read once: the entry is free again


def test_id_budget_binds(gpu, O, served):
    """With ID lists, admission is charged 4 bytes per ID + 12 per query as sized
    by each batch's count pass: a small budget defers batches, keeps the bytes
    in flight at most the budget (beyond a single batch), and changes nothing."""
    ix, q, exp = served
    ref = served_tree(O).search_dfs(q, want_ids=True)
    per_batch = [4 * int(ref["offsets"][min(i + 32, len(q))] - ref["offsets"][i]) + 12 * 32
                 for i in range(0, len(q), 32)]
    budget = int(1.5 * np.median(per_batch))  # about one batch of 32 at a time
    with gpu.Server(ix, slots=4, max_batch=32, result_budget=budget, ids=True) as sv:
        c, h, _, off, ids = sv.replay(q, np.zeros(len(q)), ids_cap=int(ref["offsets"][-1]))
        st = sv.stats()
    assert st["deferred_by_budget"] > 0, st
    assert st["bytes_peak"] <= max(budget, max(per_batch)), st
    assert_bits_equal(c, ref["counts"], "budget ID counts")
    assert_bits_equal(ids, ref["ids"], "budget ID lists")


def test_result_table_is_a_ring(gpu, served):
    """Tickets are recycled once read: a 64-entry table serves 1,000 requests
    read one by one, refuses the 65th unread submission, and forgets read ones."""
    ix, q, exp = served
    with gpu.Server(ix, max_requests=64) as sv:
        for i in range(1000):
            c, h, _ = sv.wait(sv.submit(q[i % len(q)]))
            assert c == exp["counts"][i % len(q)] and h == exp["hashes"][i % len(q)]
        tks = [sv.submit(q[i]) for i in range(64)]
        with pytest.raises(gpu.QcpError) as e:
            sv.submit(q[0])
        assert e.value.status == -6
        for i, tk in enumerate(tks):
            assert sv.wait(tk)[0] == exp["counts"][i]
        with pytest.raises(gpu.QcpError):
            sv.wait(tks[0])  # already read
        assert sv.stats()["completed"] == 1064

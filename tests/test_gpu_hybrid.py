"""CPU-GPU co-processing search (mode "hybrid"): host threads walk the upper
levels and hand the overlapping nodes of one level to the GPU.  Counts, ID-set
hashes and the canonical traversal statistics must equal the oracle's for every
hand-off level, thread count and pipeline block size."""
import numpy as np
import pytest

from helpers import assert_bits_equal, rand_boxes, rand_points, rand_queries

pytestmark = pytest.mark.gpu


def _oracle(O, objs, D, is_box, F, layout):
    return O.build_packed(objs, D, is_box, F, O.KEY_REF, layout="bvh" if layout == "bvh" else "packed")


@pytest.mark.parametrize("layout", ["mphr", "rtree", "bvh"])
@pytest.mark.parametrize("handoff", [0, 1, 2])
def test_hybrid_matches_oracle(gpu, O, layout, handoff):
    D, F = 3, 32
    is_box = layout == "rtree"
    objs = rand_boxes(60000, D, 5, mean_side=0.01) if is_box else rand_points(60000, D, 5)
    q = rand_queries(2500, D, 6, objs)
    ix = gpu.Index.build(objs, D, fanout=F, layout=layout, obj_kind="boxes" if is_box else "points")
    t = _oracle(O, objs, D, is_box, F, layout)
    exp = t.search_dfs(q)
    for threads, block in ((1, 0), (4, 700)):
        tag = "%s handoff=%d threads=%d block=%d" % (layout, handoff, threads, block)
        print("case", tag, flush=True)
        got = ix.search(q, mode="hybrid", handoff_level=handoff, cpu_threads=threads,
                        hybrid_batch=block)
        assert_bits_equal(got["counts"], exp["counts"], tag + " counts")
        assert_bits_equal(got["hashes"], exp["hashes"], tag + " hashes")
        st = got["stats"]
        assert st["internal_entries"] == int(exp["int_entries"].sum()), tag
        assert st["leaf_entries"] == int(exp["leaf_entries"].sum()), tag
        assert st["total_hits"] == int(exp["counts"].sum()), tag
        assert st["launches"] >= 1


def test_hybrid_with_containment(gpu, O):
    pts = rand_points(100000, 2, 8)
    rng = np.random.default_rng(9)
    c = rng.random((200, 2), dtype=np.float32)
    s = rng.uniform(0.01, 0.5, (200, 1)).astype(np.float32)
    q = np.concatenate([c - s / 2, c + s / 2], 1).astype(np.float32)
    ix = gpu.Index.build(pts, 2, fanout=16)
    exp = O.build_packed(pts, 2, False, 16, O.KEY_REF).search_dfs(q)
    got = ix.search(q, mode="hybrid", handoff_level=1, contain=True, cpu_threads=3)
    assert_bits_equal(got["counts"], exp["counts"], "hybrid contain counts")
    assert_bits_equal(got["hashes"], exp["hashes"], "hybrid contain hashes")


@pytest.mark.parametrize("n", [1, 40, 64, 65])
def test_hybrid_small_trees(gpu, O, n):
    pts = rand_points(n, 2, n)
    q = rand_queries(50, 2, 3, pts)
    ix = gpu.Index.build(pts, 2, fanout=64)
    exp = O.build_packed(pts, 2, False, 64, O.KEY_REF).search_dfs(q)
    got = ix.search(q, mode="hybrid")
    assert_bits_equal(got["counts"], exp["counts"], "hybrid n=%d" % n)
    assert_bits_equal(got["hashes"], exp["hashes"], "hybrid n=%d" % n)


def test_hybrid_rejects_ids_and_handles_empty(gpu):
    ix = gpu.Index.build(rand_points(1000, 2, 1), 2)
    with pytest.raises(gpu.QcpError):
        ix.search(rand_queries(10, 2, 2), mode="hybrid", ids=True)
    r = ix.search(np.zeros((0, 4), np.float32), mode="hybrid")
    assert len(r["counts"]) == 0


@pytest.mark.parametrize("layout", ["mphr", "bvh"])
def test_hybrid_tuner(gpu, O, layout):
    """The co-processing tuner times every hand-off level (the leaves as 0)
    and pipeline block that fits the sample; its choice is one of the timed
    entries, the fastest by wall time, and searches with it equal the oracle."""
    D, F = 3, 32
    pts = rand_points(80000, D, 12)
    q = rand_queries(6000, D, 13, pts)
    ix = gpu.Index.build(pts, D, fanout=F, layout=layout)
    kw, rows = gpu.autotune_hybrid(ix, q, cpu_threads=4)
    h = ix.info()["height"]
    levels = sorted({r["handoff_level"] for r in rows})
    want = list(range(1, h - 1)) if layout == "mphr" else list(range(1, min(h - 2, 6) + 1))
    assert levels == sorted([0] + want), (levels, h)
    assert {r["hybrid_batch"] for r in rows} == {4096, 16384}  # 65536 exceeds the sample
    best = min(rows, key=lambda r: r["total_ms"])
    assert (kw["handoff_level"], kw["hybrid_batch"]) == (best["handoff_level"], best["hybrid_batch"])
    exp = _oracle(O, pts, D, False, F, layout).search_dfs(q)
    got = ix.search(q, **kw)
    assert_bits_equal(got["counts"], exp["counts"], "tuned hybrid counts")
    assert_bits_equal(got["hashes"], exp["hashes"], "tuned hybrid hashes")

"""MPHR data partitions (the reference's -t, src/tree/mphr.cpp:60-108) on the
GPU against the oracle: the partitioned tree (P packed subtrees over the
reference's rank split, numbered depth-major) bit-exact in keys, permutation,
objects, topology and node MBRs; every persistent search form, the plain form
in the reference's partition mode (one launch per query, one workgroup per
partition; counts, hashes and root/node visit counters), the co-processing
form, graph-captured builds, save/load and rank-ordered builds."""
import numpy as np
import pytest

from helpers import assert_bits_equal, assert_results_equal, rand_points, rand_queries

pytestmark = pytest.mark.gpu


def _check_tree(ix, t, P, tag):
    ex = ix.export()
    assert ex["info"]["partitions"] == P, tag
    assert ex["info"]["n_nodes"] == t.n_nodes, tag
    assert_bits_equal(ex["keys"], t.keys, tag + " keys")
    assert_bits_equal(ex["perm"], t.perm, tag + " permutation")
    assert_bits_equal(ex["sobjs"], t.sobjs, tag + " rank-ordered objects")
    assert_bits_equal(ex["first"], t.first, tag + " node first")
    assert_bits_equal(ex["count"], t.count, tag + " node count")
    assert_bits_equal(ex["leaf"], t.leaf, tag + " leaf flags")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " node MBRs")


@pytest.mark.parametrize("D,P,F,n", [(2, 2, 64, 60000), (3, 7, 32, 40000), (2, 208, 64, 60000),
                                     (3, 208, 192, 50000), (2, 300, 16, 300)])
def test_partitioned_mphr(gpu, O, D, P, F, n):
    tag = "D=%d P=%d F=%d n=%d" % (D, P, F, n)
    print("case", tag, flush=True)
    pts = rand_points(n, D, seed=P * 7 + D)
    ix = gpu.Index.build(pts, D, fanout=F, layout="mphr", partitions=P)
    t = O.build_packed(pts, D, False, F, O.KEY_REF, layout="partitioned", parts=P)
    _check_tree(ix, t, P, tag)
    q = rand_queries(300, D, seed=P + 1, objs=pts)
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True)
    assert_results_equal(got, exp, tag + " persistent")
    st = got["stats"]
    assert st["internal_entries"] == int(exp["int_entries"].sum()), tag + " internal entries"
    assert st["leaf_entries"] == int(exp["leaf_entries"].sum()), tag + " leaf entries"
    for kw in ({"donate": 1, "block_waves": 8}, {"contain": True, "ids": True},
               {"order": 1, "chunks": 2, "hash": True}, {"donate": 1, "contain": True}):
        r = ix.search(q, **kw)
        assert_results_equal(r, exp, "%s %s" % (tag, kw), ids=bool(kw.get("ids")))
    # the reference's partition mode of the plain search
    mp = t.search_mprs_part(q)
    assert_bits_equal(mp["counts"], exp["counts"], tag + " oracle partition MPRS")
    for waves in (1, 3):
        pl = ix.search(q, mode="plain", visits=True, plain_waves=waves)
        assert pl["stats"]["launches"] == len(q)
        assert_bits_equal(pl["counts"], exp["counts"], tag + " plain counts")
        assert_bits_equal(pl["hashes"], exp["hashes"], tag + " plain hashes")
        assert_bits_equal(pl["visits"], mp["visits"], tag + " plain visit counters")
    # co-processing: host threads walk every partition's upper levels
    for level in (0, 1):
        hy = ix.search(q, mode="hybrid", handoff_level=level, cpu_threads=4)
        assert_results_equal(hy, exp, "%s hybrid hand-off %d" % (tag, level), ids=False)
    ix.close()


def test_partitioned_large_batch_graph_build_and_save(gpu, O, tmp_path):
    """A graph-captured build, a batch large enough to be processed in Hilbert
    order, and a save / load round trip of a partitioned index."""
    pts = rand_points(120000, 3, seed=3)
    P = 13
    t = O.build_packed(pts, 3, False, 64, O.KEY_REF, layout="partitioned", parts=P)
    ix = gpu.Index.build(pts, 3, fanout=64, partitions=P, graph=True)
    _check_tree(ix, t, P, "graph build")
    q = rand_queries(10000, 3, 4, pts)
    exp = t.search_dfs(q, want_ids=True)
    assert_results_equal(ix.search(q, ids=True), exp, "ordered batch")
    path = str(tmp_path / "part.qcp")
    ix.save(path)
    ix2 = gpu.Index.load(path)
    _check_tree(ix2, t, P, "reloaded")
    assert_results_equal(ix2.search(q[:500]), {k: v[:500] for k, v in exp.items()
                                               if k in ("counts", "hashes")},
                         "reloaded search", ids=False)
    ix.close()
    ix2.close()


def test_partitioned_ranked_and_limits(gpu, O):
    """Rank-ordered builds take partitions too (a reference dump's objects are
    re-indexed with its own partition count); P > n and non-MPHR layouts are
    refused; P = 1 is the one-tree MPHR."""
    pts = rand_points(5000, 2, seed=9)
    ix = gpu.Index.build_ranked(pts, 2, fanout=32, partitions=6)
    t = O.Tree(pts, 2, False, 32, layout="partitioned", parts=6)
    ex = ix.export()
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], "ranked partitioned MBRs")
    q = rand_queries(100, 2, 10, pts)
    assert_results_equal(ix.search(q, ids=True), t.search_dfs(q, want_ids=True), "ranked")
    with pytest.raises(gpu.QcpError) as e:
        gpu.Index.build(pts[:5], 2, fanout=32, partitions=6)
    assert e.value.status == -1
    with pytest.raises(gpu.QcpError):
        gpu.Index.build(pts, 2, fanout=32, layout="bvh", partitions=2)
    one = gpu.Index.build(pts, 2, fanout=32, partitions=1)
    plain = O.build_packed(pts, 2, False, 32, O.KEY_REF)
    assert_bits_equal(one.export()["boxes"], plain.boxes[: plain.n_nodes], "P = 1")
    assert one.info()["partitions"] == 1

"""Index from a reference MPHR dump: the objects in the dump's rank order are
indexed without re-sorting (rank IDs stay the reference's), node boxes and
search results equal the oracle's over the same rank order."""
import numpy as np
import pytest

import refdump
from helpers import assert_bits_equal, assert_results_equal, rand_points, rand_queries

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,parts,layout,F", [(3, 1, "mphr", 192), (2, 4, "mphr", 64),
                                              (3, 2, "bvh", 128)])
def test_import_reference_dump(gpu, O, tmp_path, D, parts, layout, F):
    # rank order of the dump: a shuffled order, as an unstable sort might leave ties
    pts = rand_points(50000, D, 40 + D, dup_frac=0.2)
    p = str(tmp_path / "ref.idx")
    refdump.write(p, pts, D, F=192, partitions=parts)
    objs = gpu.read_reference_mphr(p, D, 192)
    ix = gpu.Index.build_ranked(objs, D, fanout=F, layout=layout)
    t = O.Tree(objs, D, False, F, layout="bvh" if layout == "bvh" else "packed")
    ex = ix.export()
    assert_bits_equal(ex["sobjs"], objs, "rank-ordered objects")
    assert_bits_equal(ex["perm"], np.arange(len(objs), dtype=np.uint32), "identity permutation")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], "node boxes")
    q = rand_queries(300, D, 7, pts)
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True)
    assert_results_equal(got, exp, "imported %s" % layout)


@pytest.mark.parametrize("D,parts", [(2, 4), (3, 11)])
def test_import_partitioned_dump(gpu, O, tmp_path, D, parts):
    """A partitioned reference dump re-indexed with its own partition count and
    degree: every leaf of the rebuilt tree covers the rank range of a leaf of
    the dump (the same split into partitions and leaves), and the searches
    equal the oracle's partitioned tree over the dump's rank order."""
    pts = rand_points(30000, D, 60 + D)
    p = str(tmp_path / "ref_part.idx")
    refdump.write(p, pts, D, F=192, partitions=parts)
    raw = open(p, "rb").read()
    P = int(np.frombuffer(raw[:4], np.uint32)[0])
    count = int(np.frombuffer(raw[4 + 8 * P: 8 + 8 * P], np.uint32)[0])
    nodes = np.frombuffer(raw[8 + 8 * P:], refdump.node_dtype(D, 192), count=count)
    dump_leaves = sorted((int(r["index"][0]) - 1, int(r["branch_count"]))
                         for r in nodes if r["node_type"] == refdump.LEAF)
    objs = gpu.read_reference_mphr(p, D, 192)
    ix = gpu.Index.build_ranked(objs, D, fanout=192, partitions=P)
    ex = ix.export()
    ours = sorted((int(f), int(c)) for f, c, lf in zip(ex["first"], ex["count"], ex["leaf"]) if lf)
    assert ours == dump_leaves
    t = O.Tree(objs, D, False, 192, layout="partitioned", parts=P)
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], "partitioned node boxes")
    q = rand_queries(200, D, 8, pts)
    exp = t.search_dfs(q, want_ids=True)
    assert_results_equal(ix.search(q, ids=True), exp, "imported partitioned dump")
    pl = ix.search(q, mode="plain", plain_waves=3)
    assert_bits_equal(pl["counts"], exp["counts"], "imported partitioned dump, plain")

"""GPU parity of the Guttman R-tree layout (QCP_LAYOUT_GUTTMAN).

The library keys and sorts the objects on the GPU, inserts them in that order
on the host (src/tree/rtree.h quadratic split), adopts the tree in the
explicit node arrays and searches it with the persistent kernels.  Each case
compares the exported tree (keys and input indices in depth-first order,
objects, topology, exact node covers) with the oracle's Guttman build, then the
counts, ID-set hashes and sorted depth-first IDs of every search form.
"""
import struct

import numpy as np
import pytest

from helpers import (assert_bits_equal, assert_results_equal, rand_boxes, rand_points,
                     rand_queries)

pytestmark = pytest.mark.gpu


def _oracle(O, objs, D, is_box, F, key_mode="ref", key_src="lo"):
    return O.build_packed(objs, D, is_box, F, O.KEY_REF if key_mode == "ref" else O.KEY_B16,
                          O.SRC_CENTER if key_src == "center" else O.SRC_LO, layout="guttman")


def _compare_tree(ex, t, tag):
    assert ex["info"]["n"] == t.n, tag
    assert_bits_equal(ex["keys"], t.keys, tag + " keys in depth-first order")
    assert_bits_equal(ex["perm"], t.perm, tag + " input index of each position")
    assert_bits_equal(ex["sobjs"], t.sobjs, tag + " objects in depth-first order")
    assert ex["info"]["n_nodes"] == t.n_nodes, tag
    assert_bits_equal(ex["first"], t.first, tag + " node first")
    assert_bits_equal(ex["count"], t.count, tag + " node count")
    assert_bits_equal(ex["leaf"], t.leaf, tag + " leaf flags")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " exact node covers")
    assert ex["info"]["height"] == (int(t.level.max()) + 1 if t.n_nodes else 0), tag


def _check_searches(gpu, ix, t, q, tag, hybrid=True):
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True)
    assert_results_equal(got, exp, tag + " persistent")
    st = got["stats"]
    assert st["internal_entries"] == int(exp["int_entries"].sum()), tag + " internal entries"
    assert st["leaf_entries"] == int(exp["leaf_entries"].sum()), tag + " leaf entries"
    for kw in (dict(donate=0, chunks=8, order=1), dict(donate=1, block_waves=8),
               dict(hash=False, chunks=2, block_waves=1)):
        r = ix.search(q, **kw)
        if kw.get("hash", True):
            assert_results_equal(r, exp, "%s %s" % (tag, kw), ids=False)
        else:
            assert_bits_equal(r["counts"], exp["counts"], "%s %s" % (tag, kw))
    assert_results_equal(ix.search(q, ids=True, contain=True), exp, tag + " contained subtrees")
    assert_results_equal(ix.search(q, donate=1, contain=True), exp,
                         tag + " contained subtrees with donation", ids=False)
    if hybrid and t.n > 1:
        for lvl in (1, 2):
            hy = ix.search(q, mode="hybrid", handoff_level=lvl, cpu_threads=3)
            assert_results_equal(hy, exp, "%s hybrid level %d" % (tag, lvl), ids=False)
    with pytest.raises(gpu.QcpError):  # MPRS needs the packed levels
        ix.search(q, mode="plain")
    return exp


@pytest.mark.parametrize("D,is_box,F,key,src,n", [
    (2, False, 128, "ref", "lo", 60000),
    (3, False, 128, "ref", "lo", 60000),
    (2, False, 16, "b16", "lo", 30000),
    (3, False, 4, "ref", "lo", 5000),
    (2, True, 128, "ref", "center", 40000),
    (3, True, 32, "b16", "lo", 30000),
    (2, False, 256, "ref", "lo", 70000),
])
def test_guttman_parity(gpu, O, D, is_box, F, key, src, n):
    tag = "guttman D=%d box=%d F=%d key=%s src=%s n=%d" % (D, is_box, F, key, src, n)
    print("case", tag, flush=True)
    objs = rand_boxes(n, D, n + F) if is_box else rand_points(n, D, n + F, dup_frac=0.2)
    if is_box:
        objs = np.minimum(objs, np.float32(1.0))
    ix = gpu.Index.build(objs, D, fanout=F, layout="guttman", key_mode=key, key_src=src,
                         obj_kind="boxes" if is_box else "points")
    t = _oracle(O, objs, D, is_box, F, key, src)
    _compare_tree(ix.export(), t, tag)
    assert ix.info()["build_phase_ms"][4] > 0.0, tag + " host insertion time"
    _check_searches(gpu, ix, t, rand_queries(400, D, n + 1, objs), tag)
    ix.close()


@pytest.mark.parametrize("n", [0, 1, 2, 128, 129, 8192])
def test_guttman_edge_sizes(gpu, O, n):
    D, F = 2, 128
    pts = rand_points(n, D, 40 + n) if n else np.zeros((0, D), np.float32)
    print("case guttman n=%d" % n, flush=True)
    ix = gpu.Index.build(pts, D, fanout=F, layout="guttman")
    q = rand_queries(64, D, 41, pts if n else None)
    if n == 0:
        assert ix.info()["n_nodes"] == 0
        r = ix.search(q, ids=True)
        assert not r["counts"].any()
        return
    t = _oracle(O, pts, D, False, F)
    _compare_tree(ix.export(), t, "n=%d" % n)
    _check_searches(gpu, ix, t, q, "n=%d" % n)


def test_guttman_ranked_build(gpu, O):
    """qcp_build_ranked: the given order is the insertion order."""
    pts = rand_points(20000, 3, 50)
    ix = gpu.Index.build_ranked(pts, 3, fanout=64, layout="guttman")
    t = O.Tree(pts, 3, False, 64, layout="guttman")
    ex = ix.export()
    assert_bits_equal(ex["perm"], t.order, "ranked: input index of each position")
    assert_bits_equal(ex["keys"], t.order.astype(np.uint64), "ranked: keys")
    assert_bits_equal(ex["sobjs"], t.sobjs, "ranked: objects")
    assert_bits_equal(ex["first"], t.first, "ranked: first")
    assert_bits_equal(ex["count"], t.count, "ranked: count")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], "ranked: covers")
    q = rand_queries(300, 3, 51, pts)
    assert_results_equal(ix.search(q, ids=True), t.search_dfs(q, want_ids=True), "ranked search")


def test_guttman_kmeans_order(gpu, O):
    """Guttman insertion in the k-means + Hilbert order."""
    pts = rand_points(30000, 2, 60)
    init = O.kmeans_seeds(gpu.gen_words(4096, 1234, 5), len(pts), 16)
    t = O.build_kmeans(pts, 2, False, 64, O.KEY_REF, 16, 1000, init, layout="guttman")
    ix = gpu.Index.build(pts, 2, fanout=64, layout="guttman", order="kmeans", clusters=16)
    _compare_tree(ix.export(), t, "k-means guttman")
    q = rand_queries(300, 2, 61, pts)
    assert_results_equal(ix.search(q, ids=True), t.search_dfs(q, want_ids=True),
                         "k-means guttman search")


def test_guttman_rebuild_and_save_load(gpu, O, tmp_path):
    D, F, n = 2, 32, 25000
    a = rand_points(n, D, 70)
    b = rand_boxes(n, D, 71, mean_side=0.002)[:, :D].copy()  # another point set, same size
    ix = gpu.Index.build(a, D, fanout=F, layout="guttman")
    ta, tb = _oracle(O, a, D, False, F), _oracle(O, b, D, False, F)
    assert ta.n_nodes != tb.n_nodes or not np.array_equal(ta.first, tb.first)
    buf = gpu.DeviceBuffer(b)
    ix.rebuild(buf)  # the topology follows the new data
    _compare_tree(ix.export(), tb, "rebuilt")
    q = rand_queries(300, D, 72, b)
    exp = tb.search_dfs(q, want_ids=True)
    assert_results_equal(ix.search(q, ids=True), exp, "rebuilt search")
    path = str(tmp_path / "g.qcp")
    ix.save(path)
    ix2 = gpu.Index.load(path)
    _compare_tree(ix2.export(), tb, "loaded")
    assert_results_equal(ix2.search(q, ids=True), exp, "loaded search")
    ix2.close()
    # a stored topology that is not a valid tree is refused (checksum made valid)
    raw = bytearray(open(path, "rb").read())
    hdr = struct.unpack_from("<I", raw, 12)[0]
    m = ix.info()["n_nodes"]
    E = D
    count_off = hdr + 8 * n + 4 * n + 4 * E * n + 8 * D * m + 4 * m
    struct.pack_into("<I", raw, count_off, F + 1)  # root with more than F children
    h = 0xCBF29CE484222325
    for byte in bytes(raw[hdr:]):
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    struct.pack_into("<Q", raw, hdr - 8, h)
    bad = str(tmp_path / "bad.qcp")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(gpu.QcpError) as e:
        gpu.Index.load(bad)
    assert e.value.status == -5
    buf.close()
    ix.close()


def test_guttman_rejected_configurations(gpu):
    pts = rand_points(1000, 2, 80)
    for kw in (dict(fanout=3), dict(fanout=64, graph=True)):
        with pytest.raises(gpu.QcpError) as e:
            gpu.Index.build(pts, 2, layout="guttman", **kw)
        assert e.value.status == -1, kw
    ix = gpu.Index.build(pts, 2, fanout=64, layout="guttman")
    buf = gpu.DeviceBuffer(pts)
    with ix.batch(rand_queries(10, 2, 81, pts)) as b:
        with pytest.raises(gpu.QcpError):
            b.pipeline(buf)
    buf.close()
    ix.close()


def test_guttman_autotune(gpu, O):
    pts = rand_points(20000, 2, 90)
    q = rand_queries(500, 2, 91, pts)
    ix, kw, rows = gpu.autotune(pts, 2, q, fanouts=(32, 128), layout="guttman")
    F = ix.info()["fanout"]
    assert F in (32, 128) and len(rows) > 0
    t = _oracle(O, pts, 2, False, F)
    _compare_tree(ix.export(), t, "autotuned guttman F=%d" % F)
    assert_results_equal(ix.search(q, **kw), t.search_dfs(q), "autotuned guttman", ids=False)
    ix.close()

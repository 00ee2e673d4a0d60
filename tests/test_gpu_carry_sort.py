"""The object-carrying key sort (QCP_BUILD_SORT_OBJECTS): every radix pass moves
the key bits still to be sorted, the objects themselves and (unless the build
drops it) each object's input index, and the last pass writes the index's
rank-ordered objects, so no gather follows.  Every case is checked against the
CPU oracle bit for bit (keys, permutation, rank-ordered objects, topology, node
MBRs, search results) and against the index form's build of the same input:
key widths whose remainders switch from 8 to 4 bytes at every pass, zero and
one-byte keys, tile edges, more tiles than resident workgroups, every layout,
partitions, graph-captured builds and pipelines, and save/load of an index that
kept no permutation."""
import ctypes

import numpy as np
import pytest

from helpers import assert_bits_equal, assert_results_equal, rand_boxes, rand_points, rand_queries
from test_gpu_parity import check_case

pytestmark = pytest.mark.gpu


def _objs(n, D, box, seed):
    return rand_boxes(n, D, seed, mean_side=0.01) if box else rand_points(n, D, seed,
                                                                          dup_frac=0.05)


@pytest.mark.parametrize("D,box,key,src,layout,F,n", [
    (2, False, "ref", "lo", "mphr", 64, 20000),
    (3, False, "ref", "lo", "mphr", 128, 30000),
    (2, True, "ref", "center", "rtree", 128, 25000),
    (3, True, "ref", "lo", "bvh", 128, 25000),
    (3, True, "ref", "center", "rtree", 32, 25000),
    (2, False, "b16", "lo", "bvh", 16, 30000),
    (3, False, "b16", "lo", "mphr", 256, 30000),
])
@pytest.mark.parametrize("keep_perm", [True, False])
def test_carry_parity(gpu, O, D, box, key, src, layout, F, n, keep_perm):
    objs = _objs(n, D, box, seed=D * 31 + F + n % 97)
    check_case(gpu, O, objs, D, box, F, key_mode=key, key_src=src, layout=layout, nq=100,
               seed=3, sort="objects", keep_perm=keep_perm)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 8191, 8192, 8193, 16385])
def test_carry_edge_sizes(gpu, O, n):
    """Tile edges of the 8192-element sort tiles (and tiny inputs)."""
    pts = rand_points(n, 2, seed=n + 5)
    check_case(gpu, O, pts, 2, False, 64, nq=48, sort="objects")
    check_case(gpu, O, pts, 2, False, 64, layout="bvh", nq=48, sort="objects", keep_perm=False)


def test_carry_all_duplicates(gpu, O):
    """All keys equal: no pass reorders, the input order is the rank order."""
    pts = np.tile(np.array([[0.25, 0.75]], np.float32), (5000, 1))
    check_case(gpu, O, pts, 2, False, 64, nq=32, sort="objects")
    zero = np.zeros((3000, 3), np.float32)
    check_case(gpu, O, zero, 3, False, 32, nq=16, sort="objects", keep_perm=False)


@pytest.mark.parametrize("D,box,scale", [
    (2, False, 1e-5),    # 1-byte keys: the first pass is the last
    (2, False, 1e-4),    # 2 bytes
    (2, False, 1.0),     # 40 bits: 4-byte remainders from the first pass on
    (2, False, 16.0),    # 48 bits: 8-byte remainders for two passes
    (2, False, 2000.0),  # 62 bits: 8-byte remainders for four passes
    (3, False, 1.0),     # 60 bits
    (3, True, 1.0),      # 3-D boxes keyed by their lower corners (6 payload words)
])
def test_carry_key_widths(gpu, O, D, box, scale):
    n = 40000
    objs = np.ascontiguousarray(_objs(n, D, box, seed=17) * np.float32(scale), dtype=np.float32)
    layout = "rtree" if box else "mphr"
    check_case(gpu, O, objs, D, box, 64, layout=layout, nq=64, seed=9, sort="objects")
    kind = "boxes" if box else "points"
    a = gpu.Index.build(objs, D, fanout=64, layout=layout, obj_kind=kind, sort="objects").export()
    b = gpu.Index.build(objs, D, fanout=64, layout=layout, obj_kind=kind, sort="index").export()
    for k in ("keys", "perm", "sobjs", "boxes"):
        assert_bits_equal(a[k], b[k], "scale %g objects vs index form %s" % (scale, k))


def test_carry_many_tiles(gpu, O):
    """More sort tiles (611) than resident workgroups, 40-bit keys; and the
    library's default form: the index form while the objects fit the Infinity
    Cache, the object form for 2-D points beyond it (where each gathered object
    would cost a whole line)."""
    n = 5_000_000
    pts = rand_points(n, 2, seed=77)
    print("case many tiles n=%d" % n, flush=True)
    t = O.build_packed(pts, 2, False, 64, O.KEY_REF)
    for sort, keep in (("objects", True), ("objects", False)):
        ex = gpu.Index.build(pts, 2, fanout=64, sort=sort, keep_perm=keep).export()
        tag = "n=%d %s perm=%d" % (n, sort, keep)
        assert ex["info"]["sort_form"] == 1, tag
        if keep:
            assert_bits_equal(ex["perm"], t.perm, tag + " permutation")
        assert_bits_equal(ex["sobjs"], t.sobjs, tag + " rank-ordered objects")
        assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " node MBRs")
    small = gpu.Index.build(pts, 2, fanout=64)  # 40 MB: the index form
    assert small.info()["sort_form"] == 0
    small.close()
    big = rand_points(40_000_000, 2, seed=78)  # 320 MB of points: the object form
    print("case default form n=%d" % len(big), flush=True)
    ix = gpu.Index.build(big, 2, fanout=64)
    assert ix.info()["sort_form"] == 1 and ix.info()["has_perm"] == 1
    got = ix.export()
    ix.close()
    exp = gpu.Index.build(big, 2, fanout=64, sort="index").export()
    for k in ("keys", "perm", "sobjs", "boxes"):
        assert_bits_equal(got[k], exp[k], "40M points default (object) vs index form %s" % k)


def test_carry_partitions(gpu, O):
    n, D, P, F = 60000, 3, 7, 32
    pts = rand_points(n, D, seed=501)
    t = O.build_packed(pts, D, False, F, O.KEY_REF, layout="partitioned", parts=P)
    ix = gpu.Index.build(pts, D, fanout=F, partitions=P, sort="objects")
    ex = ix.export()
    for k, v in (("perm", t.perm), ("sobjs", t.sobjs), ("first", t.first), ("count", t.count),
                 ("boxes", t.boxes[: t.n_nodes])):
        assert_bits_equal(ex[k], v, "partitioned objects-form %s" % k)
    q = rand_queries(200, D, seed=502, objs=pts)
    assert_results_equal(ix.search(q, ids=True), t.search_dfs(q, want_ids=True),
                         "partitioned objects-form search")


@pytest.mark.parametrize("keep_perm", [True, False])
def test_carry_graph_key_widths(gpu, O, keep_perm):
    """One captured object-form build (every pass a 2-D REF key can need; those
    above the keys' byte width end at once, the last reordering one writes the
    index, and an all-zero input is copied in order) replayed over inputs whose
    keys need 0, 1, 2, 5, 6, 7 and 8 bytes, with eager work between replays."""
    n, D, F = 50000, 2, 64
    base = rand_points(n, D, 33)
    cases = [("zero", np.zeros_like(base)), ("x1e-5", base * np.float32(1e-5)),
             ("x1e-4", base * np.float32(1e-4)), ("x1", base), ("x16", base * np.float32(16.0)),
             ("x100", base * np.float32(100.0)), ("x2000", base * np.float32(2000.0)),
             ("x1 again", base)]
    buf = gpu.DeviceBuffer(cases[0][1])
    ix = gpu.Index.build(buf, D, fanout=F, graph=True, sort="objects", keep_perm=keep_perm)
    widths = []
    for rep, (name, objs) in enumerate(cases):
        objs = np.ascontiguousarray(objs, dtype=np.float32)
        if rep:
            buf.write(objs)
            ix.rebuild(buf)
        got = ix.export()
        assert got["info"]["sort_form"] == 1
        t = O.build_packed(objs, D, False, F, O.KEY_REF)
        widths.append((int(np.bitwise_or.reduce(t.keys)).bit_length() + 7) // 8)
        assert_bits_equal(got["keys"], t.keys, "%s keys" % name)
        assert_bits_equal(got["sobjs"], t.sobjs, "%s objects" % name)
        assert_bits_equal(got["boxes"], t.boxes, "%s boxes" % name)
        if keep_perm:
            assert_bits_equal(got["perm"], t.perm, "%s perm" % name)
        else:
            assert got["perm"] is None
        q = rand_queries(40, D, rep, objs)
        assert_results_equal(ix.search(q), t.search_dfs(q), "%s search" % name, ids=False)
    buf.close()
    assert widths[:7] == [0, 1, 2, 5, 6, 7, 8], widths


def test_carry_pipeline(gpu, O):
    """The build+query pipeline graph over an object-form index."""
    D, F = 3, 64
    a_objs, b_objs = rand_points(70000, D, 61), rand_points(70000, D, 62)
    q = rand_queries(500, D, 63, a_objs)
    exps = []
    for objs in (a_objs, b_objs):
        t = O.build_packed(objs, D, False, F, O.KEY_REF)
        exps.append(t.search_dfs(q))
    buf = gpu.DeviceBuffer(a_objs)
    ix = gpu.Index.build(buf, D, fanout=F, sort="objects", keep_perm=False)
    with ix.batch(q) as b:
        for rep in range(4):
            buf.write(a_objs if rep % 2 == 0 else b_objs)
            b.pipeline(buf)
            r = b.results()
            assert_results_equal(r, exps[rep % 2], "objects-form pipeline replay %d" % rep,
                                 ids=False)
    buf.close()


def test_carry_no_perm_export_and_save(gpu, O, tmp_path):
    n, D = 30000, 2
    pts = rand_points(n, D, seed=91)
    t = O.build_packed(pts, D, False, 64, O.KEY_REF)
    ix = gpu.Index.build(pts, D, fanout=64, sort="objects", keep_perm=False)
    assert ix.info()["has_perm"] == 0
    # a permutation request is refused, nothing else is
    perm = np.zeros(n, np.uint32)
    st = gpu.lib().qcp_export(ix._h, None, perm.ctypes.data_as(ctypes.c_void_p), None, None,
                              None, None, None)
    assert st == -1
    path = str(tmp_path / "noperm.qcp")
    ix.save(path)
    ld = gpu.Index.load(path)
    assert ld.info()["has_perm"] == 0
    ex = ld.export()
    assert ex["perm"] is None
    assert_bits_equal(ex["keys"], t.keys, "loaded keys")
    assert_bits_equal(ex["sobjs"], t.sobjs, "loaded objects")
    assert_bits_equal(ex["boxes"], t.boxes, "loaded boxes")
    q = rand_queries(200, D, 92, pts)
    assert_results_equal(ld.search(q, ids=True), t.search_dfs(q, want_ids=True), "loaded search")
    # an object-form index that kept its permutation saves and loads it
    kp = gpu.Index.build(pts, D, fanout=64, sort="objects")
    kp.save(path)
    ex2 = gpu.Index.load(path).export()
    assert_bits_equal(ex2["perm"], t.perm, "loaded permutation")

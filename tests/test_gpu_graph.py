"""hipGraph-captured build and search pipelines: every replay, with eager work
and synchronisations between replays, must equal the eager result."""
import numpy as np
import pytest

from helpers import assert_bits_equal, assert_results_equal, rand_boxes, rand_points, rand_queries

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout,D,box,F", [("mphr", 3, False, 64), ("rtree", 2, True, 128),
                                            ("bvh", 2, True, 128), ("mphr", 2, False, 32)])
def test_graph_build_replays(gpu, O, layout, D, box, F):
    mk = (lambda s: rand_boxes(60000, D, s)) if box else (lambda s: rand_points(60000, D, s))
    kind = "boxes" if box else "points"
    src = "center" if box else "lo"
    a_objs, b_objs = mk(1), mk(2)
    eager_a = gpu.Index.build(a_objs, D, fanout=F, layout=layout, obj_kind=kind, key_src=src)
    eager_b = gpu.Index.build(b_objs, D, fanout=F, layout=layout, obj_kind=kind, key_src=src)
    ea, eb = eager_a.export(), eager_b.export()
    buf = gpu.DeviceBuffer(a_objs)
    ix = gpu.Index.build(buf, D, fanout=F, layout=layout, obj_kind=kind, key_src=src, graph=True)
    for rep in range(4):
        host = a_objs if rep % 2 == 0 else b_objs
        exp = ea if rep % 2 == 0 else eb
        if rep:
            buf.write(host)
            ix.rebuild(buf)                       # replay of the captured build
        got = ix.export()
        for k in ("keys", "perm", "sobjs", "boxes", "first", "count", "leaf"):
            assert_bits_equal(got[k], exp[k], "graph build rep %d %s" % (rep, k))
        # eager work between replays
        eager_a.search(rand_queries(50, D, rep, a_objs))
    buf.close()


@pytest.mark.parametrize("donate", [0, 1])
def test_graph_search_replays(gpu, O, donate):
    pts = rand_points(200000, 3, 11)
    ix = gpu.Index.build(pts, 3, fanout=64)
    q = rand_queries(3000, 3, 12, pts)
    t = O.build_packed(pts, 3, False, 64, O.KEY_REF)
    exp = t.search_dfs(q)
    with ix.batch(q) as b:
        eager = b.run(donate=donate)
        r0 = b.results()
        assert_results_equal(r0, exp, "eager", ids=False)
        for rep in range(5):
            st = b.run(donate=donate, graph=True)
            r = b.results()
            assert_results_equal(r, exp, "graph replay %d" % rep, ids=False)
            assert st["total_hits"] == eager["total_hits"]
            assert st["leaf_entries"] == eager["leaf_entries"]
            # eager runs and other queues' work between replays
            b.run(donate=1 - donate, chunks=8)
            assert_results_equal(b.results(), exp, "eager between replays %d" % rep, ids=False)
            ix.search(q[:100], ids=True)


@pytest.mark.parametrize("layout,donate", [("mphr", 0), ("bvh", 1)])
def test_graph_build_query_pipeline(gpu, O, layout, donate):
    """Rebuild + search captured as one hipGraph: each replay (objects changed
    in place between replays) equals the eager build and search of those objects."""
    a_objs, b_objs = rand_points(80000, 3, 21), rand_points(80000, 3, 22)
    q = rand_queries(9000, 3, 23, a_objs)
    exps = [O.build_packed(o, 3, False, 64, O.KEY_REF,
                           layout="bvh" if layout == "bvh" else "packed").search_dfs(q)
            for o in (a_objs, b_objs)]
    buf = gpu.DeviceBuffer(a_objs)
    ix = gpu.Index.build(buf, 3, fanout=64, layout=layout)
    with ix.batch(q) as b:
        for rep in range(5):
            host = a_objs if rep % 2 == 0 else b_objs
            buf.write(host)
            st = b.pipeline(buf, donate=donate)
            r = b.results()
            assert_results_equal(r, exps[rep % 2], "pipeline replay %d" % rep, ids=False)
            assert st["total_hits"] == int(exps[rep % 2]["counts"].sum())
            # eager work between replays, and the rebuilt index seen from outside
            assert_bits_equal(ix.search(q[:200])["counts"], exps[rep % 2]["counts"][:200],
                              "eager search after replay %d" % rep)
    buf.close()

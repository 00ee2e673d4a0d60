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


def test_graph_build_key_widths(gpu, O):
    """A captured build runs a sort pass for every byte a key can hold (8 for
    2-D REF keys) and plans on the device: passes above the keys' actual byte
    width end at once, elements are packed from the first pass after which the
    remaining key bits and the index fit 64 bits, and the gather reads the
    result from whichever buffer the last reordering pass wrote (odd and even
    counts).  One captured graph replayed over inputs whose keys need 0, 1, 2,
    5, 6, 7 and 8 bytes (points scaled inside the REF quantiser's domain; one
    byte is the unpacked case) equals the eager build and the oracle every
    time."""
    n, D, F = 50000, 2, 64
    base = rand_points(n, D, 31)
    cases = [("zero", np.zeros_like(base)), ("x1e-5", base * np.float32(1e-5)),
             ("x1e-4", base * np.float32(1e-4)),
             ("x1", base), ("x16", base * np.float32(16.0)),
             ("x100", base * np.float32(100.0)), ("x2000", base * np.float32(2000.0)),
             ("x1 again", base)]
    widths = []
    buf = gpu.DeviceBuffer(cases[0][1])
    ix = gpu.Index.build(buf, D, fanout=F, graph=True)
    for rep, (name, objs) in enumerate(cases):
        objs = np.ascontiguousarray(objs, dtype=np.float32)
        if rep:
            buf.write(objs)
            ix.rebuild(buf)
        got = ix.export()
        t = O.build_packed(objs, D, False, F, O.KEY_REF)
        assert not t.domain_error
        widths.append((int(np.bitwise_or.reduce(t.keys)).bit_length() + 7) // 8)
        assert_bits_equal(got["keys"], t.keys, "%s keys" % name)
        assert_bits_equal(got["perm"], t.perm, "%s perm" % name)
        assert_bits_equal(got["boxes"], t.boxes, "%s boxes" % name)
        eager = gpu.Index.build(objs, D, fanout=F).export()
        for k in ("keys", "perm", "sobjs", "boxes"):
            assert_bits_equal(got[k], eager[k], "%s graph vs eager %s" % (name, k))
    buf.close()
    assert widths[:7] == [0, 1, 2, 5, 6, 7, 8], widths


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

"""Independent index handles driven from several host threads at once.  Each
handle owns its HBM tree, its scratch and its stream, and the library keeps no
mutable global state outside the allocation-free entry point, so concurrent
builds (eager and graph-captured, both sort forms), rebuilds and searches must
give exactly what the oracle gives for each handle alone."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import assert_bits_equal, assert_results_equal, rand_boxes, rand_points, rand_queries

pytestmark = pytest.mark.gpu


def _case(i):
    if i == 0:
        return rand_points(200000, 3, 11), 3, False, dict(fanout=128)
    if i == 1:
        return rand_points(300000, 2, 12), 2, False, dict(fanout=64, sort="objects", graph=True)
    if i == 2:
        return rand_boxes(150000, 2, 13), 2, True, dict(fanout=128, layout="bvh",
                                                        obj_kind="boxes")
    return rand_points(250000, 3, 14), 3, False, dict(fanout=64, layout="rtree", graph=True)


def test_concurrent_handles(gpu, O):
    cases = [_case(i) for i in range(4)]
    exps = []
    for i, (objs, D, box, kw) in enumerate(cases):
        layout = "bvh" if kw.get("layout") == "bvh" else "packed"
        t = O.build_packed(objs, D, box, kw["fanout"], O.KEY_REF, layout=layout)
        q = rand_queries(400, D, 20 + i, objs)
        exps.append((t, q, t.search_dfs(q, want_ids=True)))

    def work(i):
        objs, D, box, kw = cases[i]
        t, q, exp = exps[i]
        out = []
        buf = gpu.DeviceBuffer(objs)
        ix = gpu.Index.build(buf, D, **kw)
        for rep in range(3):
            if rep:
                ix.rebuild(buf)
            ex = ix.export()
            got = ix.search(q, ids=True)
            out.append((ex, got))
        ix.close()
        buf.close()
        return out

    with ThreadPoolExecutor(4) as pool:
        results = list(pool.map(work, range(4)))
    for i, res in enumerate(results):
        t, q, exp = exps[i]
        for rep, (ex, got) in enumerate(res):
            tag = "handle %d rebuild %d" % (i, rep)
            assert_bits_equal(ex["sobjs"], t.sobjs, tag + " rank-ordered objects")
            assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " node MBRs")
            if ex["perm"] is not None:
                assert_bits_equal(ex["perm"], t.perm, tag + " permutation")
            assert_results_equal(got, exp, tag + " search")

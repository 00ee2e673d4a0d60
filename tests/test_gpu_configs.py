"""GPU parity at the benchmark configurations, against the oracle.

Each test generates one benchmark workload (the Philox generator of qcp.py),
builds and searches it on the GPU through the C ABI, and compares with the
oracle built by its threaded build (itself checked equal to the serial oracle
in test_oracle.py):
  - the 16.8M-point 3-D mixture at full size (all 100,000 queries' counts and
    ID-set hashes in several persistent configurations, offsets of every ID
    list, and the full sorted ID lists of 1,000+ sampled queries including the
    50 heaviest);
  - the 2-D rectangles config on the packed R-tree and the BVH, searched by a
    replayed hipGraph capture with eager searches between the replays;
  - the serving config's results through the continuous-batching server;
  - the skewed 2-D config through the timed build paths (device-resident
    input, rebuild in place; eager and graph-captured) with donation on.
The tree comparisons are bit-exact: keys, permutation, rank-ordered objects,
topology and node MBRs.
"""
import os
import time

import numpy as np
import pytest

from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

C3_SCALE = 1.0     # 67,108,864 boxes and 1,000,000 queries per layout (full size)
C4_SCALE = 0.25    # 8,388,608 points, 25,000 served queries
C5_SCALE = 0.25    # 33,554,432 points, 125,000 queries


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _log(t0, what):
    print("  [%6.1f s] %s" % (time.time() - t0, what), flush=True)


def _oracle(O, objs, D, is_box, F, key_mode, key_src, layout):
    """The oracle tree of a large input: the threaded oracle build (keys, stable
    sort, gather, packed boxes), then the layout's topology and node boxes."""
    km = O.KEY_REF if key_mode == "ref" else O.KEY_B16
    ks = O.SRC_CENTER if key_src == "center" else O.SRC_LO
    keys, perm, sobjs, boxes, _, bad = O.build_mt(objs, D, is_box, F, km, ks,
                                                  nthreads=_threads())
    assert bad == 0
    t = O.Tree(sobjs, D, is_box, F, layout="bvh" if layout == "bvh" else "packed")
    if layout != "bvh":
        assert_bits_equal(t.boxes[: t.n_nodes], boxes, "threaded oracle pack")
    t.keys, t.perm = keys, perm
    return t


def _compare_tree(ix, t, tag):
    ex = ix.export()
    assert ex["info"]["n"] == t.n and ex["info"]["n_nodes"] == t.n_nodes, tag
    assert_bits_equal(ex["keys"], t.keys, tag + " sorted keys")
    if ex["perm"] is not None:  # an object-form build may keep no permutation
        assert_bits_equal(ex["perm"], t.perm, tag + " permutation")
    assert_bits_equal(ex["sobjs"], t.sobjs, tag + " rank-ordered objects")
    assert_bits_equal(ex["first"], t.first, tag + " node first")
    assert_bits_equal(ex["count"], t.count, tag + " node count")
    assert_bits_equal(ex["leaf"], t.leaf, tag + " node leaf flag")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " node MBRs")


def _sample(counts, k=1000, heavy=50, seed=0):
    """The `heavy` queries with the most hits plus k others drawn at random."""
    top = np.argsort(counts, kind="stable")[::-1][:heavy]
    rng = np.random.default_rng(seed)
    rest = rng.choice(len(counts), size=min(len(counts), k), replace=False)
    return np.unique(np.concatenate([top, rest]).astype(np.int64))


def _check_results(b, counts, hashes, tag, st=None):
    r = b.results()
    assert_bits_equal(r["counts"], counts, tag + " counts")
    assert_bits_equal(r["hashes"], hashes, tag + " hashes")
    if st is not None:
        assert st["total_hits"] == int(counts.sum(dtype=np.uint64)), tag + " total hits"


def _check_ids(b, t, queries, counts, idx, tag):
    """After an ID search of the batch: the offsets of every query's list, and the
    lists of the sampled queries, against the oracle."""
    off = b.offsets_range(0, len(queries))
    want = np.zeros(len(queries) + 1, np.uint64)
    want[1:] = np.cumsum(counts.astype(np.uint64))
    assert_bits_equal(off, want, tag + " ID list offsets")
    exp = t.search_ids_mt(queries[idx], _threads())
    assert_bits_equal(exp["counts"], counts[idx], tag + " oracle sample counts")
    for j, qi in enumerate(idx):
        o, ids = b.ids_range(int(qi), int(qi) + 1)
        e0, e1 = int(exp["offsets"][j]), int(exp["offsets"][j + 1])
        assert int(o[-1]) == e1 - e0, "%s query %d: %d IDs vs %d" % (tag, qi, int(o[-1]), e1 - e0)
        assert_bits_equal(ids, exp["ids"][e0:e1], "%s query %d IDs" % (tag, qi))
    return int(exp["offsets"][-1])


def test_c2_full_size(gpu, O):
    """16,777,216 3-D mixture points, F=128 (the tuned fanout), 100,000 queries."""
    t0 = time.time()
    w = gpu.workload("c2")
    pts, queries = w["objs"], w["queries"]
    assert len(pts) == 16777216 and len(queries) == 100000
    F = 128
    ix = gpu.Index.build(pts, 3, fanout=F, layout="mphr", key_mode="ref")
    t = _oracle(O, pts, 3, False, F, "ref", "lo", "mphr")
    _compare_tree(ix, t, "c2")
    _log(t0, "c2 tree equal (%d nodes)" % t.n_nodes)
    counts, hashes = t.search_dfs_mt(queries, _threads())
    _log(t0, "c2 oracle search: %d hits" % int(counts.sum(dtype=np.uint64)))
    with ix.batch(queries) as b:
        for kw in ({"block_waves": 8, "chunks": 8, "donate": 0}, {},
                   {"donate": 1, "chunks": 2, "block_waves": 2}, {"contain": True},
                   {"hash": True, "order": 1}):
            st = b.run(**kw)
            _check_results(b, counts, hashes, "c2 %s" % kw, st)
        b.run(ids=True, block_waves=8, chunks=8)
        idx = _sample(counts, seed=2)
        n_ids = _check_ids(b, t, queries, counts, idx, "c2")
    _log(t0, "c2 ID lists of %d sampled queries equal (%d IDs)" % (len(idx), n_ids))
    ix.close()


@pytest.mark.parametrize("layout", ["rtree", "bvh"])
def test_c3_graph_search(gpu, O, layout):
    """2-D rectangles (exponential sides, keyed by centre), F=128, the
    heavy-tailed query mix, searched by a replayed hipGraph capture."""
    t0 = time.time()
    w = gpu.workload("c3", scale=C3_SCALE)
    objs, queries = w["objs"], w["queries"]
    ix = gpu.Index.build(objs, 2, fanout=128, layout=layout, obj_kind="boxes",
                         key_src="center", graph=True)
    t = _oracle(O, objs, 2, True, 128, "ref", "center", layout)
    _compare_tree(ix, t, "c3 " + layout)
    _log(t0, "c3 %s graph-built tree equal (%d boxes)" % (layout, len(objs)))
    counts, hashes = t.search_dfs_mt(queries, _threads())
    with ix.batch(queries) as b:
        for rep in range(3):
            st = b.run(graph=True)  # captured on the first run, replayed after
            _check_results(b, counts, hashes, "c3 %s graph replay %d" % (layout, rep), st)
            st = b.run(order=1, chunks=2, donate=1)  # eager work between replays
            _check_results(b, counts, hashes, "c3 %s eager %d" % (layout, rep), st)
        st = b.run(graph=True, contain=True)
        _check_results(b, counts, hashes, "c3 %s graph contained" % layout, st)
        b.run(ids=True)
        idx = _sample(counts, seed=3)
        n_ids = _check_ids(b, t, queries, counts, idx, "c3 " + layout)
    _log(t0, "c3 %s: %d queries equal, %d sampled ID lists (%d IDs)" %
         (layout, len(queries), len(idx), n_ids))
    ix.close()


def test_c4_served_results(gpu, O):
    """Uniform 3-D points resident, Poisson arrivals at 50k QPS through the
    continuous-batching server: every served result equals the oracle's."""
    t0 = time.time()
    w = gpu.workload("c4", scale=C4_SCALE)
    pts, queries = w["objs"], w["queries"]
    ix = gpu.Index.build(pts, 3, fanout=64, layout="mphr")
    t = _oracle(O, pts, 3, False, 64, "ref", "lo", "mphr")
    _compare_tree(ix, t, "c4")
    counts, hashes = t.search_dfs_mt(queries, _threads())
    arrivals = gpu.poisson_arrivals(len(queries), 50000.0)
    with gpu.Server(ix) as sv:
        c, h, lat = sv.replay(queries, arrivals)
        stats = sv.stats()
    assert_bits_equal(c, counts, "c4 served counts")
    assert_bits_equal(h, hashes, "c4 served hashes")
    assert stats["completed"] == len(queries) and stats["batches"] >= 1
    _log(t0, "c4: %d served queries equal, p50 %.3f ms p99 %.3f ms, %d batches" %
         (len(queries), stats["p50_us"] / 1e3, stats["p99_us"] / 1e3, stats["batches"]))
    # ID-list service: every served list against the oracle's
    exp = t.search_ids_mt(queries, _threads())
    with gpu.Server(ix, ids=True, result_budget=16 << 20) as sv:
        c, h, lat, off, ids = sv.replay(queries, arrivals, ids_cap=int(exp["offsets"][-1]))
        stats = sv.stats()
    assert_bits_equal(c, counts, "c4 served ID counts")
    assert_bits_equal(off, exp["offsets"], "c4 served ID offsets")
    assert_bits_equal(ids, exp["ids"], "c4 served ID lists")
    _log(t0, "c4 ID lists: %d IDs equal, p50 %.3f ms p99 %.3f ms, peak %d result bytes" %
         (len(ids), stats["p50_us"] / 1e3, stats["p99_us"] / 1e3, stats["bytes_peak"]))
    ix.close()


def test_c5_timed_build_paths(gpu, O):
    """The skewed 2-D config from device-resident points through the timed
    build paths (build, then rebuild in place; eager and graph-captured; the
    library's form and the object-carrying sort with and without the
    permutation), searched with subtree donation on."""
    t0 = time.time()
    w = gpu.workload("c5", scale=C5_SCALE)
    pts, queries = w["objs"], w["queries"]
    t = _oracle(O, pts, 2, False, 64, "ref", "lo", "mphr")
    counts, hashes = t.search_dfs_mt(queries, _threads())
    _log(t0, "c5 oracle: %d points, %d hits" % (len(pts), int(counts.sum(dtype=np.uint64))))
    buf = gpu.DeviceBuffer(pts)
    try:
        for graph in (False, True):
            tag = "c5 %s" % ("graph" if graph else "eager")
            ix = gpu.Index.build_device(buf, 2, fanout=64, layout="mphr", graph=graph)
            ix.rebuild(buf)
            _compare_tree(ix, t, tag)
            with ix.batch(queries) as b:
                st = b.run(donate=1)
                _check_results(b, counts, hashes, tag + " donation", st)
                st = b.run(donate=1, contain=True, block_waves=8)
                _check_results(b, counts, hashes, tag + " donation + containment", st)
            if graph:
                # all lists would be 4 bytes x 1.6e11 hits: the ID search runs on a
                # batch of sampled queries (the heaviest among them)
                idx = _sample(counts, k=300, heavy=20, seed=5)
                sq = np.ascontiguousarray(queries[idx])
                with ix.batch(sq) as b:
                    b.run(ids=True, donate=1)
                    n_ids = _check_ids(b, t, sq, counts[idx], np.arange(len(idx)), tag)
                _log(t0, "%s: ID lists of %d sampled queries equal (%d IDs)" %
                     (tag, len(idx), n_ids))
            ix.close()
            _log(t0, tag + " equal")
        # the object-carrying sort (the library's form for these points beyond the
        # Infinity Cache), with and without the permutation, eager and graph
        for graph in (False, True):
            for keep in (True, False):
                tag = "c5 objects %s perm=%d" % ("graph" if graph else "eager", keep)
                ix = gpu.Index.build_device(buf, 2, fanout=64, layout="mphr", graph=graph,
                                            sort="objects", keep_perm=keep)
                ix.rebuild(buf)
                assert ix.info()["sort_form"] == 1 and ix.info()["has_perm"] == int(keep), tag
                _compare_tree(ix, t, tag)
                with ix.batch(queries) as b:
                    st = b.run(donate=1)
                    _check_results(b, counts, hashes, tag + " donation", st)
                ix.close()
                _log(t0, tag + " equal")
    finally:
        buf.close()

"""GPU parity: the HIP build and search paths against the CPU oracle, bit-exact.

Each case builds on the GPU through the C ABI (libqcp.so), exports the tree,
and compares keys, rank permutation, rank-ordered objects, node MBRs and
topology with the oracle's, then the per-query counts, ID-set hashes and sorted
rank-ID lists of the persistent form (and the plain MPRS form for packed
layouts, including its root/node visit counters).
"""
import numpy as np
import pytest

from helpers import (assert_bits_equal, assert_results_equal, rand_boxes, rand_points,
                     rand_queries)

pytestmark = pytest.mark.gpu


def _oracle_tree(O, objs, D, is_box, F, key_mode, key_src, layout):
    return O.build_packed(objs, D, is_box, F, O.KEY_REF if key_mode == "ref" else O.KEY_B16,
                          O.SRC_CENTER if key_src == "center" else O.SRC_LO,
                          layout="bvh" if layout == "bvh" else "packed")


def check_case(gpu, O, objs, D, is_box, F, key_mode="ref", key_src="lo", layout="mphr",
               nq=200, seed=0, plain=True, queries=None, sort=None, keep_perm=True,
               zero_sign=True):
    tag = "D=%d box=%d F=%d key=%s src=%s layout=%s n=%d sort=%s perm=%d" % (
        D, is_box, F, key_mode, key_src, layout, len(objs), sort, keep_perm)
    print("case", tag, flush=True)
    ix = gpu.Index.build(objs, D, fanout=F, layout=layout, key_mode=key_mode, key_src=key_src,
                         obj_kind="boxes" if is_box else "points", sort=sort,
                         keep_perm=keep_perm)
    t = _oracle_tree(O, objs, D, is_box, F, key_mode, key_src, layout)
    ex = ix.export()
    n = len(objs)
    assert ex["info"]["n"] == n
    if sort is not None and n:
        assert ex["info"]["sort_form"] == (1 if sort == "objects" else 0), tag
    assert_bits_equal(ex["keys"], t.keys, tag + " sorted keys")
    if ex["perm"] is None:
        assert not keep_perm and ex["info"]["sort_form"] == 1, tag + " permutation dropped"
    else:
        assert_bits_equal(ex["perm"], t.perm, tag + " permutation")
    assert_bits_equal(ex["sobjs"], t.sobjs, tag + " rank-ordered objects")
    assert ex["info"]["n_nodes"] == t.n_nodes, tag
    assert_bits_equal(ex["first"], t.first, tag + " node first")
    assert_bits_equal(ex["count"], t.count, tag + " node count")
    assert_bits_equal(ex["leaf"], t.leaf, tag + " node leaf flag")
    if zero_sign:
        assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], tag + " node MBRs")
    else:
        # objects holding both -0.0 and +0.0: the reference's min/max trees keep
        # whichever zero their pairing order meets first, so a zero bound's sign
        # is not pinned; every other bound is compared bit for bit
        got_b, exp_b = ex["boxes"], t.boxes[: t.n_nodes]
        assert_bits_equal(got_b == 0, exp_b == 0, tag + " node MBR zero bounds")
        assert_bits_equal(np.where(got_b == 0, 0, got_b), np.where(exp_b == 0, 0, exp_b),
                          tag + " node MBRs (zero bounds by value)")
    if queries is None:
        queries = rand_queries(nq, D, seed + 17, objs if n else None)
    exp = t.search_dfs(queries, want_ids=True)
    got = ix.search(queries, mode="persistent", ids=True)
    assert_results_equal(got, exp, tag + " persistent")
    st = got["stats"]
    assert st["internal_entries"] == int(exp["int_entries"].sum()), tag + " internal entries"
    assert st["leaf_entries"] == int(exp["leaf_entries"].sum()), tag + " leaf entries"
    assert st["total_hits"] == int(exp["counts"].sum())
    # the other persistent variants: no donation, counts only, 2 and 8 chunks in flight
    nd = ix.search(queries, mode="persistent", donate=0, chunks=8, order=1)
    assert_results_equal(nd, exp, tag + " persistent no-donation, given order", ids=False)
    co = ix.search(queries, mode="persistent", hash=False, chunks=2, block_waves=1)
    assert_bits_equal(co["counts"], exp["counts"], tag + " counts-only, 1-wave workgroups")
    w8 = ix.search(queries, mode="persistent", donate=1, block_waves=8)
    assert_results_equal(w8, exp, tag + " 8-wave workgroups with donation", ids=False)
    # containment shortcut: subtrees inside the query reported from their rank ranges
    ct = ix.search(queries, mode="persistent", ids=True, contain=True)
    assert_results_equal(ct, exp, tag + " contained subtrees")
    cd = ix.search(queries, mode="persistent", donate=1, contain=True, chunks=2)
    assert_results_equal(cd, exp, tag + " contained subtrees with donation", ids=False)
    if plain and layout != "bvh":
        mp = t.search_mprs(queries)
        assert_bits_equal(mp["counts"], exp["counts"], tag + " oracle MPRS vs DFS")
        pl = ix.search(queries, mode="plain", visits=True, plain_batch=64)
        assert_bits_equal(pl["counts"], exp["counts"], tag + " plain counts")
        assert_bits_equal(pl["hashes"], exp["hashes"], tag + " plain hashes")
        assert_bits_equal(pl["visits"], mp["visits"], tag + " plain visit counters")
    ix.close()
    return exp


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("F", [16, 32, 64, 128, 192, 256])
def test_points_ref_mphr(gpu, O, D, F):
    pts = rand_points(20000, D, seed=D * 100 + F)
    check_case(gpu, O, pts, D, False, F)


@pytest.mark.parametrize("D", [2, 3])
def test_points_b16(gpu, O, D):
    pts = rand_points(30000, D, seed=7 + D, dup_frac=0.1)
    check_case(gpu, O, pts, D, False, 64, key_mode="b16")


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("src", ["lo", "center"])
@pytest.mark.parametrize("layout", ["rtree", "bvh"])
def test_boxes(gpu, O, D, src, layout):
    objs = rand_boxes(25000, D, seed=11 * D + len(src), mean_side=0.01)
    check_case(gpu, O, objs, D, True, 128, key_src=src, layout=layout)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("F", [32, 128])
def test_points_bvh(gpu, O, D, F):
    pts = rand_points(40000, D, seed=5 + F)
    check_case(gpu, O, pts, D, False, F, layout="bvh")


@pytest.mark.parametrize("D,box,scale,n,pack_pass", [
    (2, False, 1e-5, 20000, None),   # 8-bit keys: one pass, nothing to pack
    (2, False, 1.0, 50000, 0),       # 40-bit keys: 32 + 16 bits after the first pass
    (3, False, 1.0, 50000, 1),       # 60-bit keys: 44 + 16 bits after the second
    (2, False, 2000.0, 300000, 2),   # 62-bit keys, 19-bit indices: 38 + 19 after the third
    (2, True, 1.0, 50000, 0),        # boxes keyed by their centres
    (3, True, 1.0, 50000, 1),
])
def test_packed_sort_passes(gpu, O, D, box, scale, n, pack_pass):
    """The eager build packs (key bits still to sort, input index) into one
    64-bit word from the first pass after which they fit, and its gather
    recomputes the keys from the objects; every packing point (and none) gives
    the oracle's keys, permutation, objects and node boxes."""
    objs = (rand_boxes(n, D, 41) if box else rand_points(n, D, 41)) * np.float32(scale)
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    layout = "rtree" if box else "mphr"
    t = _oracle_tree(O, objs, D, box, 64, "ref", "center" if box else "lo", layout)
    bits = int(np.bitwise_or.reduce(t.keys)).bit_length()
    passes = (bits + 7) // 8
    pbits = max(1, int(n - 1).bit_length())
    want = next((j for j in range(passes - 1) if bits - 8 * (j + 1) + pbits <= 64), None)
    assert want == pack_pass, (bits, pbits, want)
    check_case(gpu, O, objs, D, box, 64, key_src="center" if box else "lo", layout=layout,
               nq=100, seed=7)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 4095, 4096, 4097])
def test_edge_sizes(gpu, O, n):
    pts = rand_points(n, 2, seed=n)
    check_case(gpu, O, pts, 2, False, 64, nq=64)
    check_case(gpu, O, pts, 2, False, 64, layout="bvh", nq=64)


def test_all_duplicates(gpu, O):
    pts = np.tile(np.array([[0.25, 0.75]], np.float32), (5000, 1))
    check_case(gpu, O, pts, 2, False, 64, nq=32)


@pytest.mark.parametrize("layout,F", [("mphr", 16), ("mphr", 64), ("bvh", 128), ("rtree", 32)])
def test_donation_heavy_queries(gpu, O, layout, F):
    """A few very large queries and thousands of waves: idle waves take
    donated subtrees; results must not change."""
    pts = rand_points(300000, 2, seed=F)
    rng = np.random.default_rng(F)
    c = rng.random((6, 2), dtype=np.float32)
    s = np.float32(0.8)
    q = np.concatenate([c - s / 2, c + s / 2], 1).astype(np.float32)
    ix = gpu.Index.build(pts, 2, fanout=F, layout=layout)
    t = _oracle_tree(O, pts, 2, False, F, "ref", "lo", layout)
    exp = t.search_dfs(q, want_ids=True)
    for rep in range(3):
        got = ix.search(q, donate=1, ids=True)
        assert_results_equal(got, exp, "donation %s F=%d rep %d" % (layout, F, rep))
        assert got["stats"]["donated_items"] > 0
        assert got["stats"]["leaf_entries"] == int(exp["leaf_entries"].sum())
    got = ix.search(q, donate=1, contain=True)
    assert_results_equal(got, exp, "donation+contain %s F=%d" % (layout, F), ids=False)


@pytest.mark.parametrize("layout", ["mphr", "rtree", "bvh"])
@pytest.mark.parametrize("D", [2, 3])
def test_containment_large_queries(gpu, O, layout, D):
    """Queries covering large parts of the space: most hits come from subtrees
    reported whole (rank ranges, closed-form hashes, ID runs at unaligned
    output offsets); everything must equal the canonical traversal."""
    n = 200000
    pts = rand_points(n, D, seed=31 + D)
    rng = np.random.default_rng(D)
    c = rng.random((40, D), dtype=np.float32)
    s = rng.uniform(0.05, 0.9, (40, 1)).astype(np.float32)
    q = np.concatenate([c - s / 2, c + s / 2], 1).astype(np.float32)
    F = 64
    ix = gpu.Index.build(pts, D, fanout=F, layout=layout)
    t = _oracle_tree(O, pts, D, False, F, "ref", "lo", layout)
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True, contain=True)
    assert_results_equal(got, exp, "contain %s D=%d" % (layout, D))
    assert got["stats"]["contained_runs"] > 0
    assert got["stats"]["leaf_entries"] < int(exp["leaf_entries"].sum())
    dn = ix.search(q, donate=1, contain=True)
    assert_results_equal(dn, exp, "contain+donate %s D=%d" % (layout, D), ids=False)


def test_containment_off_for_inverted_boxes(gpu, O):
    """A box index holding a box with lo > hi: a node box inside the query no
    longer implies that its objects overlap, so containment is not applied."""
    objs = rand_boxes(30000, 2, seed=5, mean_side=0.01)
    objs[123, [0, 2]] = objs[123, [2, 0]] + np.float32(0.001)  # lo > hi in dimension 0
    assert objs[123, 0] > objs[123, 2]
    q = np.array([[0.0, 0.0, 1.0, 1.0], [0.1, 0.1, 0.9, 0.9]], np.float32)
    ix = gpu.Index.build(objs, 2, fanout=64, layout="rtree", obj_kind="boxes")
    t = _oracle_tree(O, objs, 2, True, 64, "ref", "lo", "rtree")
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True, contain=True)
    assert_results_equal(got, exp, "inverted box index")
    assert got["stats"]["contained_runs"] == 0


@pytest.mark.parametrize("layout", ["mphr", "bvh"])
def test_ordered_large_batch(gpu, O, layout):
    """Batches of 8192+ queries are processed in Hilbert order of their centres
    (per-XCD ranges); smaller ones as given.  Both must give the oracle's results."""
    pts = rand_points(150000, 3, seed=91)
    q = rand_queries(12000, 3, 92, pts)
    ix = gpu.Index.build(pts, 3, fanout=64, layout=layout)
    t = _oracle_tree(O, pts, 3, False, 64, "ref", "lo", layout)
    exp = t.search_dfs(q, want_ids=True)
    got = ix.search(q, ids=True)
    assert_results_equal(got, exp, "ordered batch " + layout)
    given = ix.search(q, order=1, chunks=2)
    assert_results_equal(given, exp, "given order " + layout, ids=False)
    small = ix.search(q[:3000])
    assert_bits_equal(small["hashes"], exp["hashes"][:3000], "small batch " + layout)


def test_autotune_returns_consistent_index(gpu, O):
    pts = rand_points(100000, 3, 77)
    q = rand_queries(500, 3, 78, pts)
    ix, kw, table = gpu.autotune(pts, 3, q[:200], fanouts=(16, 64))
    assert len(table) == 2 * 4 * 3 * 2
    best = min(table, key=lambda r: r["kernel_ms"])
    assert ix.info()["fanout"] == best["fanout"]
    assert kw == {"block_waves": best["block_waves"], "chunks": best["leaf_chunks"],
                  "donate": best["donate"]}
    t = O.build_packed(pts, 3, False, best["fanout"], O.KEY_REF)
    exp = t.search_dfs(q)
    got = ix.search(q, **kw)
    assert_results_equal(got, exp, "autotuned", ids=False)


def test_empty_index(gpu):
    ix = gpu.Index.build(np.zeros((0, 3), np.float32), 3)
    q = rand_queries(10, 3, 1)
    r = ix.search(q, ids=True)
    assert (r["counts"] == 0).all() and len(r["ids"]) == 0


def test_empty_query_set(gpu):
    ix = gpu.Index.build(rand_points(1000, 2, 1), 2)
    r = ix.search(np.zeros((0, 4), np.float32), ids=True)
    assert len(r["counts"]) == 0


def test_domain_error(gpu):
    pts = rand_points(100, 2, 3)
    pts[5, 1] = -0.5
    with pytest.raises(gpu.QcpError) as e:
        gpu.Index.build(pts, 2, key_mode="ref")
    assert e.value.status == -7
    pts[5, 1] = np.nan
    with pytest.raises(gpu.QcpError):
        gpu.Index.build(pts, 2, key_mode="b16")


def test_save_load_roundtrip(gpu, O, tmp_path):
    pts = rand_points(50000, 3, 9)
    for layout in ("mphr", "bvh"):
        ix = gpu.Index.build(pts, 3, fanout=64, layout=layout)
        p = str(tmp_path / ("ix_%s.qcp" % layout))
        ix.save(p)
        ix2 = gpu.Index.load(p)
        q = rand_queries(100, 3, 2, pts)
        a = ix.search(q, ids=True)
        b = ix2.search(q, ids=True)
        assert_results_equal(b, a, "reloaded " + layout)
        with open(p, "r+b") as f:
            f.seek(-5, 2)
            f.write(b"\x00\x01\x02\x03\x04")
        with pytest.raises(gpu.QcpError) as e:
            gpu.Index.load(p)
        assert e.value.status == -5


def test_c1_full_size(gpu, O):
    """C1 at full size: 1,048,576 uniform 2-D points, b16 keys, F=64,
    1,000 queries of side 0.01 -- counts and sorted ID lists bit-exact."""
    w = gpu.workload("c1")
    pts, queries = w["objs"], w["queries"]
    ix = gpu.Index.build(pts, 2, fanout=64, layout="mphr", key_mode="b16")
    t = O.build_packed(pts, 2, False, 64, O.KEY_B16)
    ex = ix.export()
    assert_bits_equal(ex["keys"], t.keys, "c1 keys")
    assert_bits_equal(ex["perm"], t.perm, "c1 perm")
    assert_bits_equal(ex["boxes"], t.boxes, "c1 node MBRs")
    exp = t.search_dfs(queries, want_ids=True)
    for mode in ("persistent",):
        got = ix.search(queries, mode=mode, ids=True)
        assert_results_equal(got, exp, "c1 " + mode)
    pl = ix.search(queries, mode="plain")
    assert_bits_equal(pl["counts"], exp["counts"], "c1 plain counts")


def test_kmeans_small_and_boxes(gpu, O):
    """Fewer objects than clusters (K becomes N), and boxes keyed by centre."""
    pts = rand_points(5, 2, 3)
    ix = gpu.Index.build(pts, 2, fanout=16, order="kmeans", clusters=64)
    init = O.kmeans_seeds(gpu.gen_words(4096, 1234, 5), 5, 5)
    t = O.build_kmeans(pts, 2, False, 16, O.KEY_REF, 5, 1000, init)
    assert_bits_equal(ix.export()["keys"], t.keys, "k-means keys, K = N")
    assert ix.info()["kmeans_iterations"] == t.iterations
    boxes = rand_boxes(30000, 2, 5, mean_side=0.01)
    init = O.kmeans_seeds(gpu.gen_words(4096, 1234, 5), len(boxes), 12)
    t = O.build_kmeans(boxes, 2, True, 64, O.KEY_REF, 12, 1000, init, key_src=O.SRC_CENTER)
    ix = gpu.Index.build(boxes, 2, fanout=64, layout="rtree", obj_kind="boxes", key_src="center",
                         order="kmeans", clusters=12)
    ex = ix.export()
    assert ex["info"]["kmeans_iterations"] == t.iterations
    assert_bits_equal(ex["keys"], t.keys, "k-means box keys")
    assert_bits_equal(ex["perm"], t.perm, "k-means box permutation")
    q = rand_queries(200, 2, 6, boxes)
    assert_results_equal(ix.search(q, ids=True), t.search_dfs(q, want_ids=True), "k-means boxes")


@pytest.mark.parametrize("D,K,layout,key,iters", [(2, 16, "mphr", "ref", 0), (3, 64, "mphr", "b16", 0),
                                                  (3, 7, "bvh", "ref", 0), (2, 300, "mphr", "ref", 4)])
def test_kmeans_hilbert_order(gpu, O, D, K, layout, key, iters):
    """k-means + Hilbert order: GPU assignment with host centre updates gives
    the oracle's clusters, composite keys, ranks, tree and search results, with
    the reference's stop rule (iters 0: at most 1001 iterations) or a cap."""
    pts = rand_points(70000, D, 80 + K)
    n = len(pts)
    init = O.kmeans_seeds(gpu.gen_words(4096, 1234, 5), n, K)
    km = O.KEY_REF if key == "ref" else O.KEY_B16
    t = O.build_kmeans(pts, D, False, 64, km, K, (iters or 1001) - 1, init,
                       layout="bvh" if layout == "bvh" else "packed")
    ix = gpu.Index.build(pts, D, fanout=64, layout=layout, key_mode=key, order="kmeans",
                         clusters=K, kmeans_iters=iters)
    ex = ix.export()
    print("k-means D=%d K=%d: %d iterations" % (D, K, t.iterations), flush=True)
    assert ex["info"]["kmeans_iterations"] == t.iterations
    if iters:
        assert t.iterations == iters
    assert_bits_equal(ex["keys"], t.keys, "k-means keys")
    assert_bits_equal(ex["perm"], t.perm, "k-means permutation")
    assert_bits_equal(ex["boxes"], t.boxes[: t.n_nodes], "k-means node boxes")
    q = rand_queries(300, D, 81, pts)
    exp = t.search_dfs(q, want_ids=True)
    assert_results_equal(ix.search(q, ids=True), exp, "k-means order search")


@pytest.mark.parametrize("n,D,K,iters", [(60000, 3, 0, 3), (1100000, 2, 40, 0),
                                         (2200000, 3, 3000, 2)])
def test_kmeans_reference_scale(gpu, O, n, D, K, iters):
    """The reference's cluster count (K 0 = 20000: twenty LDS tiles of centres,
    the last one partial) and the two- and four-objects-a-thread assignments of
    large inputs, against the oracle's clusters and composite keys."""
    pts = rand_points(n, D, 7 + D)
    Kx = min(K or 20000, n)
    init = O.kmeans_seeds(gpu.gen_words(1 << 16, 1234, 5), n, Kx)
    print("k-means n=%d D=%d K=%d iters=%d" % (n, D, Kx, iters), flush=True)
    keys, cid, its = O.kmeans_keys(pts, D, False, O.KEY_REF, Kx, (iters or 1001) - 1, init,
                                   want_iterations=True)
    ix = gpu.Index.build(pts, D, fanout=64, order="kmeans", clusters=K, kmeans_iters=iters)
    ex = ix.export()
    assert ex["info"]["kmeans_iterations"] == its
    perm = O.sort_stable(keys)
    assert_bits_equal(ex["keys"], keys[perm], "k-means keys")
    assert_bits_equal(ex["perm"], perm, "k-means permutation")
    ix.close()


def test_ids_range_pagination(gpu, O):
    pts = rand_points(50000, 2, 90)
    q = rand_queries(200, 2, 91, pts)
    ix = gpu.Index.build(pts, 2, fanout=32)
    exp = O.build_packed(pts, 2, False, 32, O.KEY_REF).search_dfs(q, want_ids=True)
    with ix.batch(q) as b:
        b.run(ids=True)
        for q0, q1 in ((0, 200), (5, 17), (199, 200), (30, 30)):
            off, ids = b.ids_range(q0, q1)
            eo = exp["offsets"][q0:q1 + 1]
            assert_bits_equal(off, eo - eo[0], "offsets %d..%d" % (q0, q1))
            assert_bits_equal(ids, exp["ids"][eo[0]:eo[-1]], "ids %d..%d" % (q0, q1))
        with pytest.raises(gpu.QcpError):
            b.ids_range(10, 201)


def test_equal_moment_id_sets(gpu, O):
    """Two queries whose hit sets {1,5,6} and {2,3,7} agree in count, sum and
    sum of squares, so their ID-set hashes are equal: every search form must
    still return the oracle's hashes, and the sorted ID lists (which the
    parity tests compare beside the hashes) tell the two sets apart."""
    x = np.arange(1, 8, dtype=np.float32) * np.float32(0.1)
    y = np.array([0.2, 0.6, 0.6, 0.9, 0.2, 0.2, 0.6], np.float32)  # rank r+1 -> y[r]
    pts = np.ascontiguousarray(np.stack([x, y], 1))
    q = np.array([[0.0, 0.1, 1.0, 0.3], [0.0, 0.5, 1.0, 0.7]], np.float32)
    t = O.Tree(pts, 2, False, 4)
    exp = t.search_dfs(q, want_ids=True)
    assert exp["ids"].tolist() == [1, 5, 6, 2, 3, 7]
    assert exp["hashes"][0] == exp["hashes"][1]
    for layout in ("mphr", "bvh"):
        ix = gpu.Index.build_ranked(pts, 2, fanout=4, layout=layout)
        assert_results_equal(ix.search(q, ids=True), exp, "equal moments " + layout)
        assert_results_equal(ix.search(q, donate=1, contain=True), exp,
                             "equal moments contained " + layout, ids=False)
        if layout == "mphr":
            pl = ix.search(q, mode="plain")
            assert_bits_equal(pl["hashes"], exp["hashes"], "equal moments plain")
        ix.close()


@pytest.mark.parametrize("layout,box", [("mphr", False), ("bvh", False), ("rtree", True)])
def test_subnormal_coordinates(gpu, O, layout, box):
    """Objects and query bounds in the FP32 subnormal range, queries whose
    bounds miss an object by one subnormal step: the leaf test's subtraction
    form equals the reference's compares only if denormals are kept."""
    rng = np.random.default_rng(17)
    n, D = 30000, 2
    tiny = np.float32(2.0 ** -149)  # the smallest subnormal
    lo = rng.integers(0, 1 << 16, size=(n, D)).astype(np.float32) * tiny
    # a tenth of the objects in the normal range, so the key sort has work
    normal = rng.random(n) < 0.1
    lo[normal] = rng.random((int(normal.sum()), D), dtype=np.float32)
    if box:
        side = rng.integers(0, 1 << 10, size=(n, D)).astype(np.float32) * tiny
        objs = np.ascontiguousarray(np.concatenate([lo, lo + side], 1), np.float32)
    else:
        objs = np.ascontiguousarray(lo)
    qs = []
    pick = rng.integers(0, n, size=64)
    for i in pick:
        p = lo[i]
        qs.append(np.concatenate([p, p]))                   # exactly on an object
        qs.append(np.concatenate([p + tiny, p + tiny]))     # one step above it
        qs.append(np.concatenate([p - tiny, p - tiny]))     # one step below it
    c = rng.integers(0, 1 << 16, size=(64, D)).astype(np.float32) * tiny
    h = rng.integers(1, 1 << 13, size=(64, 1)).astype(np.float32) * tiny
    qs.extend(np.concatenate([c - h, c + h], 1))
    q = np.ascontiguousarray(np.array(qs, np.float32))
    assert (np.abs(q[q != 0]) < np.float32(2.0 ** -126)).mean() > 0.5
    check_case(gpu, O, objs, D, box, 64, layout=layout, nq=0, plain=layout != "bvh",
               queries=q)


@pytest.mark.parametrize("D,box,layout,sort", [(2, False, "mphr", None), (2, False, "mphr", "objects"),
                                               (3, False, "mphr", None), (3, False, "bvh", "objects"),
                                               (2, True, "rtree", None), (3, True, "bvh", "objects")])
def test_special_values(gpu, O, D, box, layout, sort):
    """Infinite coordinates (b16 keys clamp them, so such objects build), signed
    zeros, and query bounds that are NaN, infinite, inverted or degenerate,
    against the reference's compares: a NaN bound rejects nothing, a query with
    lo > hi in some dimension hits nothing, and an infinite coordinate equal to
    its bound is inside.  The point test's med3 form and the box test's
    subtraction form must reproduce each case, in every search form."""
    rng = np.random.default_rng(D * 7 + int(box) * 3 + len(layout))
    n = 20000
    special = np.array([np.inf, -np.inf, 0.0, -0.0, 1.0, 0.5], np.float32)
    lo = rng.random((n, D), dtype=np.float32)
    m = rng.random((n, D)) < 0.05
    lo[m] = rng.choice(special, size=int(m.sum()))
    if box:
        side = (rng.random((n, D), dtype=np.float32) * np.float32(0.01)).astype(np.float32)
        side[(rng.random((n, D)) < 0.02) & (lo != -np.inf)] = np.inf  # boxes reaching +inf
        objs = np.ascontiguousarray(np.concatenate([lo, lo + side], 1), np.float32)
    else:
        objs = np.ascontiguousarray(lo)
    inf, nan = np.float32(np.inf), np.float32(np.nan)

    def qry(l, h):
        return np.array(list(l) + list(h), np.float32)

    qs = [qry([-inf] * D, [inf] * D),              # everything
          qry([nan] * D, [nan] * D),               # NaN bounds reject nothing
          qry([nan] + [0.25] * (D - 1), [0.5] + [nan] * (D - 1)),
          qry([inf] + [-inf] * (D - 1), [inf] * D),  # the +inf plane of dimension 0
          qry([-inf] * D, [-inf] + [inf] * (D - 1)),  # the -inf plane
          qry([0.0] * D, [0.0] * D),               # signed zeros on both bounds
          qry([-0.0] * D, [-0.0] * D),
          qry([-0.0] + [0.0] * (D - 1), [0.0] + [1.0] * (D - 1)),
          qry([0.6] + [0.0] * (D - 1), [0.4] + [1.0] * (D - 1)),  # inverted in dimension 0
          qry([0.6] + [nan] * (D - 1), [0.4] + [nan] * (D - 1)),  # inverted beside NaN bounds
          qry([1.0] * D, [1.0] * D),               # degenerate, on a special value
          qry([0.5, 0.5] + [-inf] * (D - 2), [0.5, inf] + [inf] * (D - 2))]
    for _ in range(60):  # random boxes with one bound replaced by a special value
        c = rng.random(D, dtype=np.float32)
        h = np.float32(0.05)
        ql, qh = c - h, c + h
        d = rng.integers(0, D)
        ql[d] = rng.choice(np.array([nan, -inf, -0.0, 0.0, inf], np.float32))
        d = rng.integers(0, D)
        qh[d] = rng.choice(np.array([nan, inf, 0.0, -0.0, -inf], np.float32))
        qs.append(np.concatenate([ql, qh]))
    qs.extend(rand_queries(60, D, 5, objs))
    q = np.ascontiguousarray(np.array(qs, np.float32))
    exp = check_case(gpu, O, objs, D, box, 64, key_mode="b16", layout=layout, queries=q,
                     plain=layout != "bvh", zero_sign=False, sort=sort)
    assert exp["counts"][0] == n and exp["counts"][1] == n  # the oracle agrees on the basics
    if not box:  # a box can span an inverted interval; a point cannot
        assert exp["counts"][8] == 0 and exp["counts"][9] == 0


def _max_leaf_runs(O, objs, D, F, q):
    """Most runs of consecutive leaves under one leaf-parent node of the
    oracle's packed tree whose boxes overlap query q (the reference's test)."""
    t = O.build_packed(objs, D, False, F, O.KEY_REF)
    nl = O.level_counts(len(objs), F)[-1]
    b = t.boxes[t.n_nodes - nl:t.n_nodes]
    ov = np.ones(nl, bool)
    for d in range(D):
        ov &= ~((q[d] > b[:, D + d]) | (q[D + d] < b[:, d]))
    most = 0
    for j in range(0, nl, F):
        o = ov[j:j + F].astype(np.int8)
        most = max(most, int(o[0] + ((o[1:] == 1) & (o[:-1] == 0)).sum()))
    return most


@pytest.mark.parametrize("F", [64, 128])
def test_inverted_queries_many_leaf_runs(gpu, O, F):
    """Queries with lo > hi in one dimension hit no point, but the reference
    still visits every node whose box spans the inverted interval.  Over
    uniform 3-D points a thin inverted slab gives leaf-parent nodes with ten or
    more separate runs of such leaves, more than a wave's chunks in flight:
    each run must be visited (its entries counted as visited, none scanned)
    without the traversal ever taking a leaf for an internal node."""
    n, D = 200000, 3
    objs = rand_points(n, D, seed=5)
    qs = []
    for d in range(D):
        for b in (0.3, 0.5, 0.7):
            lo = np.zeros(D, np.float32)
            hi = np.ones(D, np.float32)
            lo[d], hi[d] = np.float32(b + 1e-4), np.float32(b)
            qs.append(np.concatenate([lo, hi]))
    runs = [_max_leaf_runs(O, objs, D, F, q) for q in qs]
    assert max(runs) > 8, runs  # more runs in one node than any chunk setting keeps
    qs.extend(rand_queries(40, D, 6, objs))
    q = np.ascontiguousarray(np.array(qs, np.float32))
    exp = check_case(gpu, O, objs, D, False, F, queries=q)
    assert (exp["counts"][:len(runs)] == 0).all()
    assert (exp["leaf_entries"][:len(runs)] > 0).all()

"""Pins of the CPU oracle (no GPU): each function is checked against what the
mathematics fixes, independently of its own code, and against the committed
golden fixtures."""
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from helpers import rand_boxes, rand_points, rand_queries

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- Hilbert mapping ------------------------------------------------------------
@pytest.mark.parametrize("nd,nb", [(2, 1), (2, 2), (2, 3), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_hilbert_bijection_adjacency_roundtrip(O, nd, nb):
    side = 1 << nb
    grid = np.stack(np.meshgrid(*[np.arange(side)] * nd, indexing="ij"), -1).reshape(-1, nd)
    ks = np.array([O.c2i(nd, nb, c) for c in grid], dtype=np.uint64)
    # bijection onto [0, 2^(nd*nb))
    assert np.array_equal(np.sort(ks), np.arange(side ** nd, dtype=np.uint64))
    # consecutive keys are unit-step neighbours (the defining Hilbert property)
    g = grid[np.argsort(ks)].astype(np.int64)
    assert (np.abs(np.diff(g, axis=0)).sum(1) == 1).all()
    # the curve starts at the origin
    assert (g[0] == 0).all()
    # inverse mapping (MappingIntoMulti) round-trips
    for k in range(side ** nd):
        assert O.c2i(nd, nb, O.i2c(nd, nb, k)) == k


@pytest.mark.parametrize("nd,nb", [(2, 4), (3, 3), (2, 31), (3, 20)])
def test_hilbert_order_matches_hilbert_cmp(O, nd, nb):
    """hilbert_cmp (src/mapper/hilbert_mapper.cpp:218-262) is a separate code
    path the reference never calls; the key order must agree with it."""
    rng = np.random.default_rng(nd * 100 + nb)
    for _ in range(3000):
        a = rng.integers(0, 1 << nb, size=nd, dtype=np.uint64)
        b = rng.integers(0, 1 << nb, size=nd, dtype=np.uint64)
        if rng.random() < 0.3:   # near neighbours exercise the low digits
            b = a.copy()
            b[rng.integers(nd)] ^= np.uint64(1 << int(rng.integers(0, 3)))
        ka, kb = O.c2i(nd, nb, a), O.c2i(nd, nb, b)
        exp = -1 if ka < kb else (1 if ka > kb else 0)
        assert O.hcmp(nd, 8, nb, a, b) == exp


def test_hilbert_special_cases(O):
    # D == 1 returns the coordinate itself (src/mapper/hilbert_mapper.cpp:65-67)
    assert O.c2i(1, 20, [123456]) == 123456
    # nb == 1: no rotation walk, only the inverse-Gray prefix XOR (lines 56-62)
    for v in range(4):
        c = [v & 1, v >> 1]
        packed = c[0] | (c[1] << 1)
        g = packed
        g ^= g >> 1
        assert O.c2i(2, 1, c) == g
    # bitTranspose is the digit interleave: bit b of coordinate d -> bit b*nd+d
    rng = np.random.default_rng(3)
    for nd, nb in ((2, 31), (3, 20), (3, 7)):
        for _ in range(500):
            c = [int(x) for x in rng.integers(0, 1 << nb, size=nd)]
            packed = 0
            for d in reversed(range(nd)):
                packed = (packed << nb) | c[d]
            exp = 0
            for b in range(nb):
                for d in range(nd):
                    exp |= ((c[d] >> b) & 1) << (b * nd + d)
            assert O.lib().orc_bit_transpose(nd, nb, packed) == exp


def test_hilbert_self_similarity(O):
    """Refining the grid keeps the coarse order: the key of a cell at order nb+1
    with coordinates 2c+delta lies in the 2^nd-block of the coarse key of c
    whenever the curve's first sub-cell orientation is the identity."""
    nd, nb = 2, 4
    side = 1 << nb
    for x in range(side):
        for y in range(side):
            k0 = O.c2i(nd, nb, [x, y])
            ks = [O.c2i(nd, nb + 1, [2 * x + dx, 2 * y + dy]) for dx in (0, 1) for dy in (0, 1)]
            assert sorted(k >> 2 for k in ks) == [k0] * 4


def test_quantisers_closed_form(O):
    rng = np.random.default_rng(5)
    p = np.concatenate([rng.random(20000, dtype=np.float32),
                        np.array([0.0, 2 ** -24, 0.5, 1 - 2 ** -24, 1.0], np.float32)])
    ref = np.trunc(np.float32(1000000.0) * p).astype(np.int64)   # one FP32 multiply, truncate
    b16 = np.minimum(np.trunc(p * np.float32(65536.0)).astype(np.int64), 65535)
    for i in range(0, len(p), 37):
        assert O.lib().orc_quant_ref(float(p[i])) == ref[i]
        assert O.lib().orc_quant_b16(float(p[i])) == b16[i]
    assert O.lib().orc_quant_b16(float("nan")) == 0
    assert O.lib().orc_quant_b16(-3.0) == 0


def test_keys_domain(O):
    pts = np.array([[0.5, 0.5], [-0.1, 0.2]], np.float32)
    _, bad = O.keys(pts, 2, False, O.KEY_REF)
    assert bad == 1
    _, bad = O.keys(np.array([[0.5, 0.5, 1.5]], np.float32), 3, False, O.KEY_REF)
    assert bad == 1   # 1.5e6 >= 2^20 in 3-D ref mode
    _, bad = O.keys(np.array([[0.5, 1.5]], np.float32), 2, False, O.KEY_REF)
    assert bad == 0   # 31 bits per dimension in 2-D


def test_keys_golden(O):
    g = np.load(os.path.join(GOLD, "hilbert_keys.npz"))
    for D in (2, 3):
        for mode, name in ((O.KEY_REF, "ref"), (O.KEY_B16, "b16")):
            k, bad = O.keys(g["pts_%s_%dd" % (name, D)], D, False, mode)
            assert bad == 0
            assert np.array_equal(k, g["keys_%s_%dd" % (name, D)])
    for D, nb in ((2, 3), (3, 2)):
        grid = g["grid_%dd_%db" % (D, nb)]
        ks = np.array([O.c2i(D, nb, c) for c in grid], np.uint64)
        assert np.array_equal(ks, g["gridkeys_%dd_%db" % (D, nb)])


def test_center_key_source(O):
    objs = rand_boxes(500, 2, 1)
    kc, _ = O.keys(objs, 2, True, O.KEY_REF, O.SRC_CENTER)
    cen = ((objs[:, :2] + objs[:, 2:]) * np.float32(0.5)).astype(np.float32)
    kp, _ = O.keys(cen, 2, False, O.KEY_REF)
    assert np.array_equal(kc, kp)
    kl, _ = O.keys(objs, 2, True, O.KEY_REF, O.SRC_LO)
    kp2, _ = O.keys(np.ascontiguousarray(objs[:, :2]), 2, False, O.KEY_REF)
    assert np.array_equal(kl, kp2)


# ---- sort ------------------------------------------------------------------------
@settings(max_examples=40, deadline=None)
@given(st.lists(st.integers(0, 50), min_size=0, max_size=400))
def test_sort_stable(O, vals):
    k = np.array(vals, np.uint64)
    perm = O.sort_stable(k)
    assert np.array_equal(perm, np.lexsort((np.arange(len(k)), k)).astype(np.uint32))


# ---- packed levels -------------------------------------------------------------------
@pytest.mark.parametrize("n,F", [(2, 64), (64, 64), (65, 64), (4097, 64), (100000, 16), (1, 8)])
def test_level_counts(O, n, F):
    lv = O.level_counts(n, F)
    exp, cur = [], n
    while cur > 1:
        cur = -(-cur // F)
        exp.insert(0, cur)
    if n == 1:
        exp = [1]
    assert lv == exp


def _brute_boxes(entries_lo, entries_hi, F, sizes):
    """Padded min/max per node (thread >= count padded with 1.0 / 0.0)."""
    out_lo, out_hi, s = [], [], 0
    for c in sizes:
        lo = entries_lo[s:s + c].min(0)
        hi = entries_hi[s:s + c].max(0)
        if c < F:
            lo = np.minimum(lo, np.float32(1.0))
            hi = np.maximum(hi, np.float32(0.0))
        out_lo.append(lo)
        out_hi.append(hi)
        s += c
    return np.array(out_lo, np.float32), np.array(out_hi, np.float32)


@pytest.mark.parametrize("n,D,box,F", [(5000, 2, False, 64), (3001, 3, True, 16), (777, 2, True, 128)])
def test_pack_boxes_brute(O, n, D, box, F):
    objs = rand_boxes(n, D, n) if box else rand_points(n, D, n)
    t = O.build_packed(objs, D, box, F, O.KEY_REF)
    lv = O.level_counts(n, F)
    off = np.concatenate([[0], np.cumsum(lv)])
    lo = t.sobjs[:, :D]
    hi = t.sobjs[:, D:] if box else t.sobjs[:, :D]
    for l in range(len(lv) - 1, -1, -1):
        below = n if l == len(lv) - 1 else lv[l + 1]
        sizes = [min(F, below - j * F) for j in range(lv[l])]
        blo, bhi = _brute_boxes(lo, hi, F, sizes)
        got = t.boxes[off[l]:off[l + 1]]
        assert np.array_equal(got[:, :D].view(np.uint32), blo.view(np.uint32))
        assert np.array_equal(got[:, D:].view(np.uint32), bhi.view(np.uint32))
        lo, hi = got[:, :D], got[:, D:]


# ---- BVH topology ---------------------------------------------------------------------
def _py_split_offset(start, end):
    """GetSplitOffset by a linear scan (independent of the oracle's halving
    search): the last offset < end whose rank shares more leading bits with
    rank(start) than rank(end) does; 0 when there is none or when it is 0."""
    first, last = start + 1, end + 1
    if first >= last:
        return 0
    cp = 64 - (first ^ last).bit_length()
    best = start
    for i in range(start + 1, end):
        if 64 - (first ^ (i + 1)).bit_length() > cp:
            best = i
    return best


def _py_bvh(n, F):
    from collections import deque
    nodes = []
    q = deque([(0, n - 1, 0)])
    out = []
    while q:
        s, e, d = q.popleft()
        if e - s + 1 <= F:
            out.append(("leaf", s, e - s + 1))
            continue
        pos = [s, e]
        work = deque([(s, e)])
        while work and len(pos) < F:
            a, b = work.popleft()
            so = _py_split_offset(a, b)
            if not so:
                continue
            pos.append(so)
            if so - a >= F:
                work.append((a, so))
            if b - so + 1 >= F:
                work.append((so + 1, b))
        pos.sort()
        out.append(("int", len(pos) - 1))
        for c in range(len(pos) - 1):
            q.append((pos[c], pos[c + 1], d + 1))
            pos[c + 1] += 1
    return out


@pytest.mark.parametrize("n,F", [(129, 128), (1000, 128), (3000, 32), (700, 16)])
def test_bvh_topology_independent(O, n, F):
    t = O.Tree(np.zeros((n, 2), np.float32), 2, False, F, layout="bvh")
    ref = _py_bvh(n, F)
    assert len(ref) == t.n_nodes
    for i, r in enumerate(ref):
        if r[0] == "leaf":
            assert t.leaf[i] == 1 and t.first[i] == r[1] and t.count[i] == r[2]
        else:
            assert t.leaf[i] == 0 and t.count[i] == r[1]


@pytest.mark.parametrize("n", [129, 1000, 16384, 100000])
def test_bvh_topology_invariants_and_golden(O, n):
    g = np.load(os.path.join(GOLD, "bvh_topology.npz"))
    t = O.Tree(np.zeros((n, 2), np.float32), 2, False, 128, layout="bvh")
    assert np.array_equal(t.first, g["first_%d" % n])
    assert np.array_equal(t.count, g["count_%d" % n])
    assert np.array_equal(t.leaf, g["leaf_%d" % n])
    leaves = np.nonzero(t.leaf)[0]
    # leaves partition the ranks in BFS order... in rank order after sorting by first
    f = t.first[leaves]
    c = t.count[leaves]
    order = np.argsort(f)
    assert f[order][0] == 0 and (f[order][1:] == (f + c)[order][:-1]).all()
    assert (f + c).max() == n and (c <= 128).all() and (c >= 1).all()
    internal = np.nonzero(t.leaf == 0)[0]
    assert (t.count[internal] <= 127).all()
    # children of each internal node are consecutive ids after it
    for i in internal:
        assert t.first[i] > i


# ---- searches ------------------------------------------------------------------------------
@pytest.mark.parametrize("n,D,box,F,layout", [
    (6000, 2, False, 64, "packed"), (5000, 3, False, 16, "packed"), (4000, 2, True, 128, "packed"),
    (7000, 3, True, 128, "bvh"), (9000, 2, False, 32, "bvh"), (1, 2, False, 64, "packed"),
    (65, 3, False, 64, "bvh"), (200, 2, False, 2, "packed")])
def test_search_dfs_mprs_brute_agree(O, n, D, box, F, layout):
    objs = rand_boxes(n, D, n) if box else rand_points(n, D, n, dup_frac=0.05)
    t = O.build_packed(objs, D, box, F, O.KEY_REF, layout=layout)
    q = rand_queries(300, D, n + 1, objs)
    a = t.search_dfs(q, want_ids=True)
    b = O.search_brute(t.sobjs, D, box, q, want_ids=True)
    for k in ("counts", "hashes", "offsets", "ids"):
        assert np.array_equal(a[k], b[k]), k
    if layout == "packed":
        c = t.search_mprs(q, want_ids=True)
        for k in ("counts", "hashes", "offsets", "ids"):
            assert np.array_equal(c[k], b[k]), k
    # per-query ID lists are strictly ascending
    for i in range(len(q)):
        ids = b["ids"][b["offsets"][i]:b["offsets"][i + 1]]
        assert (np.diff(ids.astype(np.int64)) > 0).all()
    # hash closed form in exact integers: (sum of squares + K * sum) mod 2^64
    K = 0x9E3779B97F4A7C15
    for i in range(0, len(q), 7):
        ids = [int(x) for x in b["ids"][b["offsets"][i]:b["offsets"][i + 1]]]
        want = (sum(x * x for x in ids) + K * sum(ids)) % (1 << 64)
        assert int(b["hashes"][i]) == want
        assert O.id_hash(b["ids"][b["offsets"][i]:b["offsets"][i + 1]]).sum(dtype=np.uint64) == want


def test_search_numpy_closed_form(O):
    """Brute force against a numpy closed-interval filter (independent code)."""
    objs = rand_boxes(3000, 3, 4)
    q = rand_queries(100, 3, 5, objs)
    r = O.search_brute(objs, 3, True, q, want_ids=True)
    for i in range(len(q)):
        lo, hi = q[i, :3], q[i, 3:]
        ok = ~((lo > objs[:, 3:]) | (hi < objs[:, :3])).any(1)
        exp = np.nonzero(ok)[0].astype(np.uint32) + 1
        assert np.array_equal(r["ids"][r["offsets"][i]:r["offsets"][i + 1]], exp)


@settings(max_examples=25, deadline=None)
@given(st.integers(1, 3000), st.sampled_from([2, 3]), st.sampled_from([4, 16, 64, 128]),
       st.integers(0, 10 ** 6), st.booleans())
def test_search_property(O, n, D, F, seed, bvh):
    objs = rand_points(n, D, seed, dup_frac=0.1)
    t = O.build_packed(objs, D, False, F, O.KEY_REF, layout="bvh" if bvh else "packed")
    q = rand_queries(40, D, seed + 1, objs)
    a = t.search_dfs(q)
    b = O.search_brute(t.sobjs, D, False, q)
    assert np.array_equal(a["counts"], b["counts"]) and np.array_equal(a["hashes"], b["hashes"])


def test_search_golden(O):
    g = np.load(os.path.join(GOLD, "search_boxes3d.npz"))
    for layout, F in (("packed", 16), ("bvh", 32), ("packed", 128)):
        t = O.build_packed(g["objs"], 3, True, F, O.KEY_REF, O.SRC_CENTER, layout=layout)
        assert np.array_equal(t.perm, g["perm"])
        r = t.search_dfs(g["queries"], want_ids=True)
        for k in ("counts", "hashes", "offsets", "ids"):
            assert np.array_equal(r[k], g[k]), (layout, k)


def test_dfs_mt_matches_serial(O):
    objs = rand_points(20000, 3, 8)
    t = O.build_packed(objs, 3, False, 64, O.KEY_REF)
    q = rand_queries(257, 3, 9, objs)
    a = t.search_dfs(q)
    c, h = t.search_dfs_mt(q, 5)
    assert np.array_equal(a["counts"], c) and np.array_equal(a["hashes"], h)


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("D,is_box", [(2, False), (3, False), (2, True)])
def test_threaded_cpu_build_equals_serial(O, threads, D, is_box):
    """The threaded CPU build (the CPU baseline) gives the serial oracle's keys,
    permutation, rank-ordered objects and node boxes, duplicates included."""
    objs = rand_boxes(20011, D, 3 + D, mean_side=0.01) if is_box else rand_points(20011, D, 3 + D, dup_frac=0.2)
    k, perm, sobjs, boxes, ms, bad = O.build_mt(objs, D, is_box, 64, O.KEY_REF, nthreads=threads)
    t = O.build_packed(objs, D, is_box, 64, O.KEY_REF)
    assert bad == 0
    assert np.array_equal(k, t.keys) and np.array_equal(perm, t.perm)
    assert np.array_equal(sobjs.view(np.uint32), t.sobjs.view(np.uint32))
    assert np.array_equal(boxes.view(np.uint32), t.boxes.view(np.uint32))
    assert (ms >= 0).all()


def test_threaded_mprs_equals_serial(O):
    pts = rand_points(30000, 3, 21)
    q = rand_queries(300, 3, 22, pts)
    t = O.build_packed(pts, 3, False, 32, O.KEY_REF)
    ref = t.search_mprs(q)
    c, h = t.search_mprs_mt(q, 5)
    assert np.array_equal(c, ref["counts"]) and np.array_equal(h, ref["hashes"])


def _kmeans_numpy(pts, K, max_loop, init):
    """Independent numpy restatement of the k-means rule: FP32 distances summed
    over d in order, first minimum wins; FP64 means summed sequentially within
    chunks of 65536, chunks in order, after every assignment; repeat while
    float32(changes) / float32(n) > 1e-5 and loop++ < max_loop (the reference's
    condition, src/mapper/kmeans_mapper.cpp:170)."""
    n, D = pts.shape
    cen = pts[init].copy()
    cid = np.full(n, 0xFFFFFFFF, np.uint64)
    loop = its = 0
    while True:
        dd = np.zeros((n, K), np.float32)
        for d in range(D):
            t = pts[:, d:d + 1] - cen[None, :, d]
            dd = dd + t * t
        new = np.argmin(dd, axis=1).astype(np.uint64)
        changes = int((new != cid).sum())
        cid = new
        its += 1
        for k in range(K):
            tot, s = 0, np.zeros(D)
            for c0 in range(0, n, 65536):
                sel = pts[c0:c0 + 65536][cid[c0:c0 + 65536] == k].astype(np.float64)
                tot += len(sel)
                if len(sel):
                    s = s + np.cumsum(sel, axis=0)[-1]
            if tot:
                cen[k] = (s / tot).astype(np.float32)
        delta = np.float32(changes) / np.float32(n)
        if not (float(delta) > 1e-5):
            break
        if not loop < max_loop:
            break
        loop += 1
    return cid.astype(np.uint32), cen, its


@pytest.mark.parametrize("D,K,max_loop", [(2, 5, 1000), (3, 8, 1000), (2, 16, 2), (3, 7, 0)])
def test_kmeans_order_against_numpy(O, D, K, max_loop):
    n = 3000 if K < 16 else 70000
    pts = rand_points(n, D, 60 + D)
    init = np.arange(0, n, n // K, dtype=np.uint32)[:K]
    keys, cid, its = O.kmeans_keys(pts, D, False, O.KEY_REF, K, max_loop, init,
                                   want_iterations=True)
    cid_np, cen, its_np = _kmeans_numpy(pts, K, max_loop, init)
    assert its == its_np
    assert its <= max_loop + 1
    if max_loop == 1000:
        assert its < 1000  # converged by the delta rule, not the cap
    assert np.array_equal(cid, cid_np)
    # composite key: cluster rank (by centre key) above the object's key
    obj_keys, _ = O.keys(pts, D, False, O.KEY_REF)
    nb = 31 if D == 2 else 20
    ck = np.array([O.c2i(D, nb, [int(np.float32(1e6) * np.float32(c)) for c in cen[k]])
                   for k in range(K)], np.uint64)
    rank = np.argsort(np.argsort(ck, kind="stable"), kind="stable")
    kbits = int(np.ceil(np.log2(K)))
    s = max(0, D * nb + kbits - 64)
    want = (rank[cid].astype(np.uint64) << np.uint64(D * nb - s)) | (obj_keys >> np.uint64(s))
    assert np.array_equal(keys, want)
    # the sorted order groups clusters by rank, Hilbert order inside
    perm = O.sort_stable(keys)
    assert (np.diff(rank[cid[perm]].astype(np.int64)) >= 0).all()


def test_kmeans_seeds_rule(O, qcp):
    w = qcp.gen_words(64, 1234, 5)
    s = O.kmeans_seeds(w, 1000, 10)
    assert len(set(s.tolist())) == 10
    assert s[0] == (int(w[0]) * 1000) >> 32


def test_id_hash_collides_on_equal_moment_sets(O):
    """The ID-set hash depends only on the count, the sum and the sum of squares
    of the IDs, so {1,5,6} and {2,3,7} collide; the GPU tests therefore compare
    sorted ID lists wherever they compare hashes of small sets.  Also the per-ID
    C term against exact integers."""
    a, b = np.array([1, 5, 6], np.uint32), np.array([2, 3, 7], np.uint32)
    ha = sum(O.lib().orc_id_hash(int(x)) for x in a) % (1 << 64)
    hb = sum(O.lib().orc_id_hash(int(x)) for x in b) % (1 << 64)
    assert ha == hb
    assert int(O.id_hash(a).sum(dtype=np.uint64)) == ha
    for x in (1, 2, 3, 1000, 65535, 65536, 2 ** 31 - 1, 2 ** 32 - 513):
        want = (x * x + O.HASH_K * x) % (1 << 64)
        assert O.lib().orc_id_hash(x) == want
        assert int(O.id_hash(np.array([x], np.uint32))[0]) == want


@pytest.mark.parametrize("n,F,P", [(1000, 16, 3), (5000, 64, 7), (60000, 64, 208), (300, 32, 300),
                                   (4097, 64, 2), (129, 128, 1)])
def test_partitioned_mphr(O, n, F, P):
    """MPHR data partitions: the reference's split (chunk = n/P, the first
    partition also takes n%P), and every partition's subtree equal to the
    packed tree of its rank range built on its own (orc_packed_topology +
    orc_pack), found by a BFS from root p of the depth-major numbering; then
    DFS over the whole tree and the per-partition MPRS against brute force."""
    sizes = [O.partition_range(n, P, p) for p in range(P)]
    assert sizes[0] == (0, n // P + n % P)
    assert all(c == n // P for _, c in sizes[1:])
    assert all(sizes[p + 1][0] == sizes[p][0] + sizes[p][1] for p in range(P - 1))
    assert sizes[-1][0] + sizes[-1][1] == n
    objs = rand_points(n, 2, seed=n + P)
    k, _ = O.keys(objs, 2, False, O.KEY_REF)
    sobjs = np.ascontiguousarray(objs[O.sort_stable(k)])
    t = O.Tree(sobjs, 2, False, F, layout="partitioned", parts=P)
    depth_of = np.zeros(t.n_nodes, np.int64)
    for p, (b, c) in enumerate(sizes):
        sub = O.Tree(sobjs[b:b + c], 2, False, F)  # the partition alone, packed
        bfs, i = [p], 0
        while i < len(bfs):
            v = bfs[i]
            i += 1
            if not t.leaf[v]:
                bfs.extend(range(int(t.first[v]), int(t.first[v]) + int(t.count[v])))
        bfs = np.array(bfs)
        assert len(bfs) == sub.n_nodes
        assert np.array_equal(t.count[bfs], sub.count) and np.array_equal(t.leaf[bfs], sub.leaf)
        leaves = sub.leaf.astype(bool)
        assert np.array_equal(t.first[bfs][leaves], sub.first[leaves] + b)
        assert np.array_equal(t.boxes[bfs].view(np.uint32), sub.boxes[: sub.n_nodes].view(np.uint32))
        depth_of[bfs] = sub.level
    # depth-major numbering: depths never decrease with the node id, roots are 0..P-1
    assert (np.diff(depth_of) >= 0).all() and (depth_of[:P] == 0).all()
    q = rand_queries(120, 2, seed=P, objs=sobjs)
    brute = O.search_brute(sobjs, 2, False, q, want_ids=True)
    dfs = t.search_dfs(q, want_ids=True)
    for key in ("counts", "hashes", "offsets", "ids"):
        assert np.array_equal(dfs[key], brute[key]), key
    mp = t.search_mprs_part(q)
    assert np.array_equal(mp["counts"], brute["counts"])
    assert np.array_equal(mp["hashes"], brute["hashes"])
    assert (mp["visits"][:, 0] >= P).all()  # every partition's root is visited at least once
    if P == 1:
        ref = O.Tree(sobjs, 2, False, F).search_mprs(q)
        assert np.array_equal(mp["visits"], ref["visits"])


def test_partitioned_refuses_empty_partitions(O):
    nf = np.zeros(64, np.uint32)
    assert O.lib().orc_partitioned_topology(5, 4, 6, 64, O.ptr(nf), O.ptr(nf), O.ptr(nf), O.ptr(nf)) == -1

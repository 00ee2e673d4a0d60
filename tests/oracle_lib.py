"""ctypes bindings of the CPU oracle (oracle/_build/libqcp_oracle.so).

Test infrastructure: imported by the tests, the smoke check and the
CPU-baseline leg of bench.py only.  The oracle restates the reference's CPU
path; see oracle/qcp_oracle.h for the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libqcp_oracle.so")

KEY_REF, KEY_B16 = 0, 1
HASH_K = 0x9E3779B97F4A7C15
SRC_LO, SRC_CENTER = 0, 1

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I64 = ctypes.c_int64
INT = ctypes.c_int
UNS = ctypes.c_uint


def _declare(L):
    L.orc_bit_transpose.argtypes = [UNS, UNS, U64]
    L.orc_bit_transpose.restype = U64
    L.orc_hilbert_c2i.argtypes = [UNS, UNS, P]
    L.orc_hilbert_c2i.restype = U64
    L.orc_hilbert_i2c.argtypes = [UNS, UNS, U64, P]
    L.orc_hilbert_cmp.argtypes = [UNS, UNS, UNS, P, P]
    L.orc_hilbert_cmp.restype = INT
    L.orc_quant_ref.argtypes = [ctypes.c_float]
    L.orc_quant_ref.restype = I64
    L.orc_quant_b16.argtypes = [ctypes.c_float]
    L.orc_quant_b16.restype = U32
    L.orc_keys.argtypes = [P, U64, UNS, INT, INT, INT, P]
    L.orc_keys.restype = INT
    L.orc_sort_stable.argtypes = [P, U64, P]
    L.orc_level_counts.argtypes = [U64, UNS, P, INT]
    L.orc_level_counts.restype = INT
    L.orc_pack.argtypes = [P, U64, UNS, INT, UNS, P]
    L.orc_pack.restype = U64
    L.orc_bvh_topology.argtypes = [U64, UNS, I64, P, P, P, P]
    L.orc_bvh_topology.restype = I64
    L.orc_packed_topology.argtypes = [U64, UNS, I64, P, P, P, P]
    L.orc_packed_topology.restype = I64
    L.orc_tree_boxes.argtypes = [P, UNS, INT, UNS, I64, P, P, P, P]
    L.orc_id_hash.argtypes = [U32]
    L.orc_id_hash.restype = U64
    L.orc_search_dfs.argtypes = [P, UNS, INT, I64, P, P, P, P, P, U64, U64,
                                 P, P, P, P, P, P]
    L.orc_search_dfs_mt.argtypes = [P, UNS, INT, I64, P, P, P, P, P, U64, UNS, P, P]
    L.orc_search_mprs.argtypes = [P, U64, UNS, INT, UNS, P, P, U64, P, P, P, P, P]
    L.orc_search_brute.argtypes = [P, U64, UNS, INT, P, U64, P, P, P, P]
    L.orc_build_mt.argtypes = [P, U64, UNS, INT, INT, INT, UNS, UNS, P, P, P, P, P]
    L.orc_build_mt.restype = INT
    L.orc_search_mprs_mt.argtypes = [P, U64, UNS, INT, UNS, P, P, U64, UNS, P, P]
    L.orc_kmeans_keys.argtypes = [P, U64, UNS, INT, INT, INT, UNS, UNS, P, P, P, P]
    L.orc_guttman_build.argtypes = [P, U64, UNS, INT, UNS, I64, P, P, P, P, P]
    L.orc_guttman_build.restype = I64
    L.orc_partition_range.argtypes = [U64, UNS, UNS, P, P]
    L.orc_search_dfs_roots.argtypes = [P, UNS, INT, I64, P, P, P, P, UNS, P, U64, U64,
                                       P, P, P, P, P, P]
    L.orc_search_dfs_mt_roots.argtypes = [P, UNS, INT, I64, P, P, P, P, UNS, P, U64, UNS, P, P]
    L.orc_partitioned_topology.argtypes = [U64, UNS, UNS, I64, P, P, P, P]
    L.orc_partitioned_topology.restype = I64
    L.orc_search_mprs_part.argtypes = [P, U64, UNS, INT, UNS, UNS, P, U64, P, P, P]


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def c2i(nd, nb, coords):
    c = np.ascontiguousarray(coords, dtype=np.uint64)
    return lib().orc_hilbert_c2i(nd, nb, ptr(c))


def i2c(nd, nb, index):
    c = np.zeros(nd, dtype=np.uint64)
    lib().orc_hilbert_i2c(nd, nb, index, ptr(c))
    return c


def hcmp(nd, nbytes, nb, c1, c2):
    a = np.ascontiguousarray(c1, dtype=np.dtype("<u%d" % nbytes))
    b = np.ascontiguousarray(c2, dtype=np.dtype("<u%d" % nbytes))
    return lib().orc_hilbert_cmp(nd, nbytes, nb, ptr(a), ptr(b))


def keys(objs, D, is_box, key_mode, key_src=SRC_LO):
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    n = objs.shape[0]
    out = np.zeros(n, dtype=np.uint64)
    bad = lib().orc_keys(ptr(objs), n, D, int(is_box), key_mode, key_src, ptr(out))
    return out, bad


def sort_stable(k):
    k = np.ascontiguousarray(k, dtype=np.uint64)
    perm = np.zeros(len(k), dtype=np.uint32)
    lib().orc_sort_stable(ptr(k), len(k), ptr(perm))
    return perm


def level_counts(n, F):
    out = np.zeros(64, dtype=np.uint64)
    h = lib().orc_level_counts(n, F, ptr(out), 64)
    return [int(x) for x in out[:h]]


def guttman(objs, D, is_box, F):
    """Guttman insertion of objs in index order: (first, count, leaf, level, order)."""
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    n = objs.shape[0]
    cap = max(16, 2 * n + 16)
    nf = np.zeros(cap, dtype=np.uint32)
    nc = np.zeros(cap, dtype=np.uint32)
    nl = np.zeros(cap, dtype=np.uint8)
    lv = np.zeros(cap, dtype=np.uint8)
    order = np.zeros(max(n, 1), dtype=np.uint32)
    m = lib().orc_guttman_build(ptr(objs), n, D, int(bool(is_box)), F, cap, ptr(nf), ptr(nc),
                                ptr(nl), ptr(lv), ptr(order))
    assert m >= 0
    return nf[:m].copy(), nc[:m].copy(), nl[:m].copy(), lv[:m].copy(), order[:n].copy()


class Tree:
    """An explicit tree (BFS order) over rank-sorted objects, as the oracle builds it.
    layout "guttman" inserts the rank-sorted objects one by one and then holds
    them in depth-first leaf order (self.order maps position -> rank)."""

    def __init__(self, sobjs, D, is_box, F, layout="packed", parts=1):
        self.sobjs = np.ascontiguousarray(sobjs, dtype=np.float32)
        self.n = self.sobjs.shape[0]
        self.D, self.is_box, self.F, self.layout = D, int(bool(is_box)), F, layout
        self.parts = parts
        cap = max(16, 2 * self.n + 16)
        nf = np.zeros(cap, dtype=np.uint32)
        nc = np.zeros(cap, dtype=np.uint32)
        nl = np.zeros(cap, dtype=np.uint8)
        lv = np.zeros(cap, dtype=np.uint8)
        L = lib()
        if layout == "guttman":
            f, c, lf, lvl, self.order = guttman(self.sobjs, D, is_box, F)
            m = len(f)
            nf[:m], nc[:m], nl[:m], lv[:m] = f, c, lf, lvl
            self.sobjs = np.ascontiguousarray(self.sobjs[self.order])
        elif layout == "bvh":
            m = L.orc_bvh_topology(self.n, F, cap, ptr(nf), ptr(nc), ptr(nl), ptr(lv))
        elif layout == "partitioned":
            m = L.orc_partitioned_topology(self.n, F, parts, cap, ptr(nf), ptr(nc), ptr(nl),
                                           ptr(lv))
        else:
            m = L.orc_packed_topology(self.n, F, cap, ptr(nf), ptr(nc), ptr(nl), ptr(lv))
        assert m >= 0
        self.n_nodes = m
        self.first, self.count, self.leaf, self.level = nf[:m].copy(), nc[:m].copy(), nl[:m].copy(), lv[:m].copy()
        self.boxes = np.zeros((max(m, 1), 2 * D), dtype=np.float32)
        if layout in ("bvh", "guttman", "partitioned"):
            L.orc_tree_boxes(ptr(self.sobjs), D, self.is_box, 0 if layout == "guttman" else F, m,
                             ptr(self.first), ptr(self.count), ptr(self.leaf), ptr(self.boxes))
        else:
            got = L.orc_pack(ptr(self.sobjs), self.n, D, self.is_box, F, ptr(self.boxes))
            assert got == m

    def _args(self):
        return (ptr(self.sobjs), self.D, self.is_box, self.n_nodes, ptr(self.first),
                ptr(self.count), ptr(self.leaf), ptr(self.boxes),
                getattr(self, "parts", 1) if getattr(self, "layout", "") == "partitioned" else 1)

    def search_dfs(self, queries, want_ids=False):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts = np.zeros(nq, dtype=np.uint32)
        hashes = np.zeros(nq, dtype=np.uint64)
        ie = np.zeros(nq, dtype=np.uint64)
        le = np.zeros(nq, dtype=np.uint64)
        lib().orc_search_dfs_roots(*self._args(), ptr(q), 0, nq, ptr(counts), ptr(hashes),
                             None, None, ptr(ie), ptr(le))
        res = {"counts": counts, "hashes": hashes, "int_entries": ie, "leaf_entries": le}
        if want_ids:
            offsets = np.zeros(nq + 1, dtype=np.uint64)
            offsets[1:] = np.cumsum(counts.astype(np.uint64))
            ids = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
            lib().orc_search_dfs_roots(*self._args(), ptr(q), 0, nq, None, None,
                                 ptr(offsets), ptr(ids), None, None)
            res["offsets"] = offsets
            res["ids"] = ids[: int(offsets[-1])]
        return res

    def search_dfs_mt(self, queries, nthreads):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts = np.zeros(nq, dtype=np.uint32)
        hashes = np.zeros(nq, dtype=np.uint64)
        lib().orc_search_dfs_mt_roots(*self._args(), ptr(q), nq, nthreads, ptr(counts), ptr(hashes))
        return counts, hashes

    def search_ids_mt(self, queries, nthreads):
        """Counts, hashes, offsets and sorted ID lists of every query, the DFS run
        over nthreads Python threads (contiguous query ranges; the C oracle runs
        without the GIL)."""
        from concurrent.futures import ThreadPoolExecutor

        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts, hashes = self.search_dfs_mt(q, nthreads)
        offsets = np.zeros(nq + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(counts.astype(np.uint64))
        ids = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
        parts = 8 * nthreads  # more ranges than threads: heavy queries cluster
        cuts = [nq * t // parts for t in range(parts + 1)]
        L = lib()

        def run(t):
            if cuts[t] < cuts[t + 1]:
                L.orc_search_dfs_roots(*self._args(), ptr(q), cuts[t], cuts[t + 1], None, None,
                                 ptr(offsets), ptr(ids), None, None)

        with ThreadPoolExecutor(max_workers=nthreads) as ex:
            list(ex.map(run, range(parts)))
        return {"counts": counts, "hashes": hashes, "offsets": offsets,
                "ids": ids[: int(offsets[-1])]}

    def search_mprs_mt(self, queries, nthreads):
        assert self.layout == "packed"
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts = np.zeros(nq, dtype=np.uint32)
        hashes = np.zeros(nq, dtype=np.uint64)
        lib().orc_search_mprs_mt(ptr(self.sobjs), self.n, self.D, self.is_box, self.F,
                                 ptr(self.boxes), ptr(q), nq, nthreads, ptr(counts), ptr(hashes))
        return counts, hashes

    def search_mprs_part(self, queries):
        """The plain search over every partition (per-query sums of counts,
        hashes, root and node visits)."""
        assert self.layout == "partitioned"
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts = np.zeros(max(nq, 1), dtype=np.uint32)
        hashes = np.zeros(max(nq, 1), dtype=np.uint64)
        visits = np.zeros(2 * max(nq, 1), dtype=np.uint32)
        lib().orc_search_mprs_part(ptr(self.sobjs), self.n, self.D, self.is_box, self.F,
                                   self.parts, ptr(q), nq, ptr(counts), ptr(hashes), ptr(visits))
        return {"counts": counts[:nq], "hashes": hashes[:nq], "visits": visits[: 2 * nq].reshape(nq, 2)}

    def search_mprs(self, queries, want_ids=False):
        assert self.layout == "packed"
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        counts = np.zeros(nq, dtype=np.uint32)
        hashes = np.zeros(nq, dtype=np.uint64)
        visits = np.zeros(2 * nq, dtype=np.uint32)
        lib().orc_search_mprs(ptr(self.sobjs), self.n, self.D, self.is_box, self.F,
                              ptr(self.boxes), ptr(q), nq, ptr(counts), ptr(hashes),
                              None, None, ptr(visits))
        res = {"counts": counts, "hashes": hashes, "visits": visits.reshape(nq, 2)}
        if want_ids:
            offsets = np.zeros(nq + 1, dtype=np.uint64)
            offsets[1:] = np.cumsum(counts.astype(np.uint64))
            ids = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
            lib().orc_search_mprs(ptr(self.sobjs), self.n, self.D, self.is_box, self.F,
                                  ptr(self.boxes), ptr(q), nq, None, None, ptr(offsets),
                                  ptr(ids), None)
            res["offsets"] = offsets
            res["ids"] = ids[: int(offsets[-1])]
        return res


def search_brute(sobjs, D, is_box, queries, want_ids=False):
    s = np.ascontiguousarray(sobjs, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = q.shape[0]
    counts = np.zeros(nq, dtype=np.uint32)
    hashes = np.zeros(nq, dtype=np.uint64)
    lib().orc_search_brute(ptr(s), s.shape[0], D, int(is_box), ptr(q), nq, ptr(counts),
                           ptr(hashes), None, None)
    res = {"counts": counts, "hashes": hashes}
    if want_ids:
        offsets = np.zeros(nq + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(counts.astype(np.uint64))
        ids = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
        lib().orc_search_brute(ptr(s), s.shape[0], D, int(is_box), ptr(q), nq, None, None,
                               ptr(offsets), ptr(ids))
        res["offsets"] = offsets
        res["ids"] = ids[: int(offsets[-1])]
    return res


def id_hash(ids):
    """numpy restatement of orc_id_hash per ID (uint64 wraparound): id^2 + K*id."""
    x = ids.astype(np.uint64)
    with np.errstate(over="ignore"):
        return x * x + np.uint64(HASH_K) * x


def build_mt(objs, D, is_box, F, key_mode, key_src=SRC_LO, nthreads=8):
    """The threaded CPU build (keys, parallel stable sort, gather, packed boxes);
    returns (keys, perm, sobjs, boxes, phase ms, domain error)."""
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    n = objs.shape[0]
    E = 2 * D if is_box else D
    keys = np.zeros(max(n, 1), np.uint64)
    perm = np.zeros(max(n, 1), np.uint32)
    sobjs = np.zeros((max(n, 1), E), np.float32)
    m = sum(level_counts(n, F)) if n else 0
    boxes = np.zeros((max(m, 1), 2 * D), np.float32)
    ms = np.zeros(4, np.float64)
    bad = lib().orc_build_mt(ptr(objs), n, D, int(is_box), key_mode, key_src, F, nthreads,
                             ptr(keys), ptr(perm), ptr(sobjs), ptr(boxes), ptr(ms))
    return keys[:n], perm[:n], sobjs[:n], boxes[:m], ms, bad


def kmeans_seeds(words, n, K):
    """Seed objects of the k-means order: (w * n) >> 32 over the Philox words in
    order, repeats skipped."""
    out, seen = [], set()
    for w in words:
        i = (int(w) * n) >> 32
        if i in seen:
            continue
        seen.add(i)
        out.append(i)
        if len(out) == K:
            return np.array(out, np.uint32)
    raise ValueError("not enough words")


def kmeans_keys(objs, D, is_box, key_mode, K, max_loop, init_idx, key_src=SRC_LO,
                want_iterations=False):
    """Composite sort keys and clusters of the k-means order (max_loop: the
    reference's loop limit, 1000 there)."""
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    n = objs.shape[0]
    keys = np.zeros(max(n, 1), np.uint64)
    cid = np.zeros(max(n, 1), np.uint32)
    its = ctypes.c_uint()
    init = np.ascontiguousarray(init_idx, np.uint32)
    lib().orc_kmeans_keys(ptr(objs), n, D, int(is_box), key_mode, key_src, K, max_loop, ptr(init),
                          ptr(keys), ptr(cid), ctypes.byref(its))
    if want_iterations:
        return keys[:n], cid[:n], its.value
    return keys[:n], cid[:n]


def build_kmeans(objs, D, is_box, F, key_mode, K, max_loop, init_idx, key_src=SRC_LO,
                 layout="packed"):
    """Oracle build in the k-means + Hilbert order (max_loop as in kmeans_keys)."""
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    k, cid, its = kmeans_keys(objs, D, is_box, key_mode, K, max_loop, init_idx, key_src,
                              want_iterations=True)
    perm = sort_stable(k)
    t = Tree(objs[perm], D, is_box, F, layout=layout)
    if layout == "guttman":
        perm = perm[t.order]
    t.keys = k[perm]
    t.perm = perm
    t.cid = cid
    t.iterations = its
    return t


def partition_range(n, P, p):
    b, c = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_partition_range(n, P, p, ctypes.byref(b), ctypes.byref(c))
    return b.value, c.value


def build_packed(objs, D, is_box, F, key_mode, key_src=SRC_LO, layout="packed", parts=1):
    """Full oracle build: keys -> stable sort -> packed, BVH, Guttman or
    partitioned MPHR tree."""
    objs = np.ascontiguousarray(objs, dtype=np.float32)
    k, bad = keys(objs, D, is_box, key_mode, key_src)
    perm = sort_stable(k)
    sobjs = objs[perm]
    t = Tree(sobjs, D, is_box, F, layout=layout, parts=parts)
    if layout == "guttman":
        perm = perm[t.order]
    t.keys = k[perm]
    t.perm = perm
    t.domain_error = bad
    return t

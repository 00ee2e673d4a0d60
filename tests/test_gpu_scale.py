"""An index over more than 2^31 objects: 2,147,487,745 uniform 2-D points
(17.2 GB), built on one GPU in both key-sort forms, so that every rank, input
index and packed sort element needs the 32nd bit.  The oracle cannot build this
size in a test's time, so the checks are the properties that pin the build:
- the sorted keys never decrease, and equal keys keep their input order;
- the permutation is a permutation, and each rank-ordered object is the input
  object it names;
- sampled keys equal the oracle's keys of their objects;
- sampled leaf boxes are the min/max of their 64 objects, and the root box
  covers everything;
- three searches equal a numpy scan of all points, as sorted input indices;
- the object form and the index form give bit-identical indexes.
"""
import ctypes
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

N = (1 << 31) + 4097
F = 64
CHUNK = 1 << 26
_T0 = time.time()


def _say(msg):
    print("scale n=%d, %.0f s: %s" % (N, time.time() - _T0, msg), flush=True)


def _chunks():
    return [(s, min(N, s + CHUNK)) for s in range(0, N, CHUNK)]


def _par(fn, items):
    """fn over items on 16 threads (numpy's array loops release the GIL)."""
    with ThreadPoolExecutor(16) as ex:
        return list(ex.map(fn, items))


def _export(gpu, ix, with_keys):
    inf = ix.info()
    n, m = inf["n"], inf["n_nodes"]
    keys = np.empty(n, np.uint64) if with_keys else None
    perm = np.empty(n, np.uint32)
    sobjs = np.empty((n, 2), np.float32)
    boxes = np.empty((m, 4), np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    st = gpu.lib().qcp_export(ix._h, p(keys), p(perm), p(sobjs), p(boxes), None, None, None)
    assert st == 0, "qcp_export status %d" % st
    return keys, perm, sobjs, boxes


def test_more_than_2_31_objects(gpu, O):
    global _T0
    _T0 = time.time()
    rng = np.random.default_rng(2031)
    _say("generating points")
    objs = np.empty((N, 2), np.float32)
    seeds = np.random.SeedSequence(2031).spawn(len(_chunks()))

    def gen(i):
        s, e = _chunks()[i]
        objs[s:e] = np.random.default_rng(seeds[i]).random((e - s, 2), dtype=np.float32)

    _par(gen, range(len(seeds)))
    buf = gpu.DeviceBuffer(objs)
    _say("building (object form)")
    ix = gpu.Index.build(buf, 2, fanout=F, sort="objects")
    inf = ix.info()
    assert inf["n"] == N and inf["sort_form"] == 1 and inf["has_perm"] == 1
    _say("built in %.1f ms; exporting" % inf["build_ms"])
    keys, perm, sobjs, boxes = _export(gpu, ix, True)

    _say("keys sorted, ties in input order")

    def sorted_chunk(c):
        s, e = c[0], min(N - 1, c[1])
        a, b = keys[s:e], keys[s + 1:e + 1]
        ok = bool((a <= b).all())
        tie = a == b
        return ok and bool((perm[s:e][tie] < perm[s + 1:e + 1][tie]).all())

    assert all(_par(sorted_chunk, _chunks())), "keys decrease or a tie leaves input order"

    _say("permutation and rank-ordered objects")
    seen = np.zeros(N, np.bool_)

    def mark(c):
        seen[perm[c[0]:c[1]]] = True  # every writer stores True

    _par(mark, _chunks())
    assert all(_par(lambda c: bool(seen[c[0]:c[1]].all()), _chunks())), "not a permutation"
    del seen

    def same(c):
        s, e = c
        return np.array_equal(sobjs[s:e].view(np.uint32), objs[perm[s:e]].view(np.uint32))

    assert all(_par(same, _chunks())), "a rank-ordered object differs from its input object"

    _say("sampled keys against the oracle")
    ranks = np.concatenate([rng.integers(0, N, 200000), np.arange(N - 5000, N)])
    assert_bits_equal(keys[ranks],
                      O.keys(np.ascontiguousarray(sobjs[ranks]), 2, False, O.KEY_REF)[0],
                      "sampled keys")

    _say("leaf and root boxes")
    nleaf = (N + F - 1) // F
    leaf0 = len(boxes) - nleaf  # packed layout: leaves are the last level
    js = np.concatenate([rng.integers(0, nleaf - 1, 20000), [0, nleaf - 2, nleaf - 1]])
    for j in js:
        seg = sobjs[j * F:min(N, (j + 1) * F)]
        lo, hi = seg.min(0), seg.max(0)
        if len(seg) < F:  # the reference's padding of a partial node: lower 1.0, upper 0.0
            lo, hi = np.minimum(lo, np.float32(1.0)), np.maximum(hi, np.float32(0.0))
        assert_bits_equal(boxes[leaf0 + j], np.concatenate([lo, hi]), "leaf %d box" % j)
    mm = np.array(_par(lambda c: np.concatenate([objs[c[0]:c[1]].min(0), objs[c[0]:c[1]].max(0)]),
                       _chunks()))
    root = np.concatenate([mm[:, :2].min(0), mm[:, 2:].max(0)]).astype(np.float32)
    assert_bits_equal(boxes[0], root, "root box")

    _say("searches against a scan of all points")
    q = np.array([[0.25, 0.25, 0.25 + 1e-4, 0.25 + 2e-4],
                  [0.9, 0.1, 0.9 + 3e-3, 0.1 + 1e-3],
                  [0.0, 0.0, 2e-3, 1.0]], np.float32)
    got = ix.search(q, ids=True)
    off = np.concatenate([[0], np.cumsum(got["counts"].astype(np.int64))])
    for i in range(len(q)):
        def scan(c):
            o = objs[c[0]:c[1]]
            m = (o[:, 0] >= q[i, 0]) & (o[:, 0] <= q[i, 2]) & (o[:, 1] >= q[i, 1]) & (o[:, 1] <= q[i, 3])
            return np.nonzero(m)[0] + c[0]

        exp = np.concatenate(_par(scan, _chunks()))
        ids = got["ids"][off[i]:off[i + 1]].astype(np.int64)
        assert got["counts"][i] == len(exp), "query %d count" % i
        assert_bits_equal(np.sort(perm[ids - 1]).astype(np.int64), exp, "query %d inputs" % i)
    assert got["counts"][2] > 4_000_000
    ix.close()

    _say("index form")
    ix2 = gpu.Index.build(buf, 2, fanout=F, sort="index")
    assert ix2.info()["sort_form"] == 0
    _say("built in %.1f ms; exporting" % ix2.info()["build_ms"])
    _, perm2, sobjs2, boxes2 = _export(gpu, ix2, False)
    ix2.close()
    buf.close()
    assert_bits_equal(perm2, perm, "index form permutation")
    assert_bits_equal(sobjs2, sobjs, "index form objects")
    assert_bits_equal(boxes2, boxes, "index form node boxes")
    _say("done")

"""qcp_search_noalloc, the search entry point that allocates nothing (its
scratch is a fixed array inside the library), against the persistent search
of qcp_batch_run and the oracle: two calls in a row per configuration, the
second with fewer queries, bit-exact counts and ID-set hashes; no change in
the device's free memory across calls; sizes beyond the fixed array refused."""
import ctypes

import numpy as np
import pytest

from helpers import assert_bits_equal, rand_points, rand_queries

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _run(gpu, ix, bufs, q, **kw):
    qb, cb, hb = bufs
    hash_ = kw.get("hash", True)
    gpu.search_noalloc(ix, qb, q, cb, hb if hash_ else None, **kw)
    out = {"counts": cb.read()[:q]}
    if hash_:
        out["hashes"] = hb.read()[:q]
    return out


@pytest.mark.parametrize("layout", ["mphr", "bvh"])
def test_noalloc_equals_batch_run(gpu, O, layout):
    pts = rand_points(300000, 3, seed=5)
    ix = gpu.Index.build(pts, 3, fanout=64, layout=layout)
    t = O.build_packed(pts, 3, False, 64, O.KEY_REF, layout="bvh" if layout == "bvh" else "packed")
    q1 = rand_queries(20000, 3, 6, pts)  # 8192 or more: processed in Hilbert order
    q2 = rand_queries(3000, 3, 7, pts)   # the second, smaller call: as given
    n = len(q1)
    qb = gpu.DeviceBuffer(q1)
    cb = gpu.DeviceBuffer(np.zeros(n, np.uint32))
    hb = gpu.DeviceBuffer(np.zeros(n, np.uint64))
    try:
        for kw in ({}, {"donate": 1}, {"hash": False}, {"contain": True, "block_waves": 8,
                                                        "chunks": 8, "donate": 1}):
            for qs in (q1, q2):
                exp = ix.search(qs, **kw)
                buf = np.zeros((n, 6), np.float32)
                buf[: len(qs)] = qs
                qb.write(buf)
                got = _run(gpu, ix, (qb, cb, hb), len(qs), **kw)
                tag = "noalloc %s %s q=%d" % (layout, kw, len(qs))
                assert_bits_equal(got["counts"], exp["counts"], tag + " counts")
                if "hashes" in got:
                    assert_bits_equal(got["hashes"], exp["hashes"], tag + " hashes")
                    ora = t.search_dfs(qs)
                    assert_bits_equal(got["hashes"], ora["hashes"], tag + " oracle hashes")
        # repeated calls allocate nothing: free device memory is unchanged
        qb.write(q1)
        _run(gpu, ix, (qb, cb, hb), n, donate=1)
        before = _free_bytes()
        for _ in range(3):
            _run(gpu, ix, (qb, cb, hb), n, donate=1)
            _run(gpu, ix, (qb, cb, hb), len(q2))
        assert _free_bytes() == before
        # beyond the fixed array, and a missing hash buffer: refused
        with pytest.raises(gpu.QcpError) as e:
            gpu.search_noalloc(ix, qb, gpu.NOALLOC_MAX_QUERIES + 1, cb, hb)
        assert e.value.status == -1
        with pytest.raises(gpu.QcpError):
            gpu.search_noalloc(ix, qb, n, cb, None, hash=True)
    finally:
        for b in (qb, cb, hb):
            b.close()
        ix.close()

"""Writes the committed golden fixtures under tests/golden/ from the CPU oracle.

The reference ships no test vectors, data files or expected outputs, so these
fixtures are the oracle's outputs on small, fully specified inputs.  Each one is
also pinned independently in tests/ (Hilbert invariants, closed-form
quantisers, brute-force boxes and searches, an independent BVH split rule), so
they guard against regressions of the oracle and give the GPU tests fixed
expected values.

Run: python tests/golden/make_golden.py   (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def hilbert_fixture():
    out = {}
    rng = np.random.default_rng(20240601)
    for D in (2, 3):
        for mode, name in ((O.KEY_REF, "ref"), (O.KEY_B16, "b16")):
            edge = np.array([0.0, 2.0 ** -24, 0.5, 1.0 - 2.0 ** -24, 1e-6, 0.999999, 0.25,
                             0.125], np.float32)
            pts = rng.random((4000, D), dtype=np.float32)
            grid = np.stack(np.meshgrid(*[edge] * D, indexing="ij"), -1).reshape(-1, D)
            pts = np.concatenate([pts, grid]).astype(np.float32)
            if mode == O.KEY_B16:
                pts = np.concatenate([pts, np.ones((1, D), np.float32)])
            k, bad = O.keys(pts, D, False, mode)
            assert bad == 0
            out["pts_%s_%dd" % (name, D)] = pts
            out["keys_%s_%dd" % (name, D)] = k
    for D, nb in ((2, 3), (3, 2)):
        side = 1 << nb
        grid = np.stack(np.meshgrid(*[np.arange(side)] * D, indexing="ij"), -1).reshape(-1, D)
        out["grid_%dd_%db" % (D, nb)] = grid.astype(np.uint64)
        out["gridkeys_%dd_%db" % (D, nb)] = np.array([O.c2i(D, nb, c) for c in grid], np.uint64)
    np.savez_compressed(os.path.join(HERE, "hilbert_keys.npz"), **out)


def bvh_fixture():
    out = {}
    for n in (129, 1000, 16384, 100000):
        t = O.Tree(np.zeros((n, 2), np.float32), 2, False, 128, layout="bvh")
        out["first_%d" % n] = t.first
        out["count_%d" % n] = t.count
        out["leaf_%d" % n] = t.leaf
    np.savez_compressed(os.path.join(HERE, "bvh_topology.npz"), **out)


def search_fixture():
    """Small 3-D box config: objects, queries, per-query counts, hashes and
    sorted rank-ID lists (BVH F=32 and packed F=16 must both give them)."""
    rng = np.random.default_rng(77)
    n = 3000
    lo = rng.random((n, 3), dtype=np.float32)
    objs = np.concatenate([lo, lo + rng.exponential(0.02, (n, 3)).astype(np.float32)], 1)
    objs = objs.astype(np.float32)
    c = rng.random((64, 3), dtype=np.float32)
    s = (10.0 ** rng.uniform(-2, -0.5, (64, 1))).astype(np.float32)
    q = np.concatenate([c - np.float32(0.5) * s, c + np.float32(0.5) * s], 1).astype(np.float32)
    t = O.build_packed(objs, 3, True, 16, O.KEY_REF, O.SRC_CENTER)
    r = t.search_dfs(q, want_ids=True)
    np.savez_compressed(os.path.join(HERE, "search_boxes3d.npz"), objs=objs, queries=q,
                        perm=t.perm, counts=r["counts"], hashes=r["hashes"],
                        offsets=r["offsets"], ids=r["ids"])


def c1_fixture():
    """C1 expected outputs: 1,048,576 uniform 2-D points (Philox stream 0,
    seed 1234), b16 keys, F=64, 1,000 queries of side 0.01."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(HERE))
    spec = importlib.util.spec_from_file_location(
        "qcp", os.path.join(root, "query-co-processing-on-heterogeneous-architectures_amd", "qcp.py"))
    qcp = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(qcp)
    w = qcp.workload("c1")
    t = O.build_packed(w["objs"], 2, False, 64, O.KEY_B16)
    r = t.search_dfs(w["queries"], want_ids=True)
    k = 25
    np.savez_compressed(os.path.join(HERE, "c1_expected.npz"), counts=r["counts"],
                        hashes=r["hashes"], ids_head=r["ids"][: int(r["offsets"][k])],
                        offsets_head=r["offsets"][: k + 1], keys_head=t.keys[:4096],
                        perm_head=t.perm[:4096], root_box=t.boxes[0],
                        data_head=w["objs"][:16], query_head=w["queries"][:16])


if __name__ == "__main__":
    hilbert_fixture()
    bvh_fixture()
    search_fixture()
    c1_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))

"""Randomised parity sweep: seeded random trees (dimension, object kind, layout,
fanout, key mode, size, MPHR partitions) and random search configurations
(workgroup size, leaf chunks, donation, containment, query order, mode), each
compared with the CPU oracle's DFS results (counts, ID-set hashes, and sorted ID
lists where asked).  Fanouts that are not multiples of 4 start leaf runs off
the 16-byte boundary the 4-entries-a-lane leaf path needs."""
import numpy as np
import pytest

from helpers import assert_bits_equal, assert_results_equal, rand_boxes, rand_points, rand_queries

pytestmark = pytest.mark.gpu

N_CASES = 40


def _case(rng):
    D = int(rng.choice([2, 3]))
    layout = str(rng.choice(["mphr", "rtree", "bvh"]))
    is_box = bool(rng.random() < 0.4)
    F = int(rng.choice([16, 24, 30, 32, 64, 66, 100, 128, 200, 256])) if layout != "bvh" else \
        int(rng.choice([16, 32, 64, 128]))
    key = "b16" if rng.random() < 0.25 else "ref"
    n = int(rng.choice([1, 2, 17, 300, 4097, 20000, 60000]))
    return D, layout, is_box, F, key, n


@pytest.mark.parametrize("case", range(N_CASES))
def test_random_configuration(gpu, O, case):
    rng = np.random.default_rng(1000 + case)
    D, layout, is_box, F, key, n = _case(rng)
    src = "center" if is_box and rng.random() < 0.5 else "lo"
    seed = int(rng.integers(1 << 30))
    objs = rand_boxes(n, D, seed, mean_side=float(rng.choice([0.001, 0.02]))) if is_box \
        else rand_points(n, D, seed, dup_frac=float(rng.choice([0.0, 0.3])))
    if is_box:  # keep every corner inside the 3-D quantiser's domain [0, 1.048576)
        objs = np.minimum(objs, np.float32(1.0))
    nq = int(rng.choice([1, 50, 400]))
    q = rand_queries(nq, D, seed + 1, objs)
    if rng.random() < 0.3:  # some large queries (containment, long leaf runs)
        c = rng.random((nq, D), dtype=np.float32)
        s = rng.uniform(0.05, 0.8, (nq, 1)).astype(np.float32)
        q = np.concatenate([c - s / 2, c + s / 2], 1).astype(np.float32)
    P = int(rng.choice([2, 5])) if layout == "mphr" and n >= 5 and rng.random() < 0.5 else 1
    tag = "case %d: D=%d %s box=%d F=%d key=%s src=%s n=%d nq=%d P=%d" % (
        case, D, layout, is_box, F, key, src, n, nq, P)
    print(tag, flush=True)
    ix = gpu.Index.build(objs, D, fanout=F, layout=layout, key_mode=key, key_src=src,
                         obj_kind="boxes" if is_box else "points", partitions=P)
    t = O.build_packed(objs, D, is_box, F, O.KEY_B16 if key == "b16" else O.KEY_REF,
                       O.SRC_CENTER if src == "center" else O.SRC_LO,
                       layout="bvh" if layout == "bvh" else ("partitioned" if P > 1 else "packed"),
                       parts=P)
    exp = t.search_dfs(q, want_ids=True)
    for rep in range(3):
        kw = dict(block_waves=int(rng.choice([1, 2, 4, 8])), chunks=int(rng.choice([2, 4, 8])),
                  donate=int(rng.random() < 0.4), contain=bool(rng.random() < 0.4),
                  order=int(rng.random() < 0.3))
        ids = bool(rng.random() < 0.5) and not kw["donate"]
        got = ix.search(q, ids=ids, **kw)
        assert_results_equal(got, exp, "%s %s ids=%d" % (tag, kw, ids), ids=ids)
    if layout != "bvh" and rng.random() < 0.5:
        pl = ix.search(q, mode="plain", plain_batch=int(rng.choice([1, 7, 128])))
        assert_bits_equal(pl["counts"], exp["counts"], tag + " plain")
    if n > 1 and rng.random() < 0.5:
        hy = ix.search(q, mode="hybrid", handoff_level=int(rng.choice([0, 1, 2])),
                       cpu_threads=int(rng.choice([1, 3])), hybrid_batch=int(rng.choice([0, 37])))
        assert_bits_equal(hy["counts"], exp["counts"], tag + " hybrid")
        assert_bits_equal(hy["hashes"], exp["hashes"], tag + " hybrid hashes")

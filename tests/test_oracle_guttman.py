"""The oracle's Guttman R-tree (src/tree/rtree.h insertion, quadratic split,
spherical-volume metric) pinned on the CPU.

Pins, chosen so that a plausible slip fails one of them:
  - hand-derived splits and descents at node capacity 4 (minimum fill 2):
    the seeds must be the pair wasting the most volume, entries must join in
    order of decreasing growth difference, the minimum fill must force the
    last entry into the other group, the first group must stay in the old
    node, and the descent must take the entry that grows least;
  - structural invariants at larger sizes: one leaf level, F/2..F entries per
    non-root node, depth-first positions tiling [0, N), the insertion order a
    permutation, and exact (unpadded) covers;
  - search: DFS over the tree equals the brute-force filter of the objects in
    depth-first order.
"""
import numpy as np
import pytest

from helpers import assert_bits_equal, rand_boxes, rand_points, rand_queries

LEAF = 0x80000000


def _tree(O, objs, D, is_box, F):
    return O.guttman(np.asarray(objs, dtype=np.float32), D, is_box, F)


def test_split_hand_derived(O):
    # A, B, C, D cluster near (0.1, 0.1); E is far.  Capacity 4, so E splits
    # the root leaf over the buffer [A, B, C, D, E]:
    #  - seeds A, E (the pair whose bounding sphere is largest: all are points);
    #  - growth differences toward {A} vs {E}: D 0.296*pi, B 0.280*pi,
    #    C 0.240*pi, so D joins {A} first, then B ({A,D} vs {E}: B 0.280*pi
    #    against C 0.240*pi), and {A,D,B} reaches total - fill = 3;
    #  - the minimum fill sends C to {E};
    #  - group 0 stays in the old node in buffer order [A, B, D]; [C, E] is the
    #    new sibling; the new root holds [old, new].
    pts = np.array([[0.10, 0.10], [0.20, 0.10], [0.10, 0.30], [0.13, 0.13], [0.90, 0.90]],
                   dtype=np.float32)
    first, count, leaf, level, order = _tree(O, pts, 2, False, 4)
    assert list(leaf) == [0, 1, 1]
    assert list(count) == [2, 3, 2]
    assert list(first) == [1, 0, 3]
    assert list(level) == [0, 1, 1]
    assert list(order) == [0, 1, 3, 2, 4]


def test_descent_hand_derived(O):
    # After the split above: leaf0 covers [0.1,0.2]x[0.1,0.13], leaf1 covers
    # [0.1,0.9]x[0.3,0.9].  (0.5, 0.5) lies inside leaf1's cover (growth 0)
    # and goes there; (0.11, 0.12) lies inside leaf0's (growth 0).
    pts = np.array([[0.10, 0.10], [0.20, 0.10], [0.10, 0.30], [0.13, 0.13], [0.90, 0.90],
                    [0.50, 0.50], [0.11, 0.12]], dtype=np.float32)
    first, count, leaf, level, order = _tree(O, pts, 2, False, 4)
    assert list(count) == [2, 4, 3]
    assert list(order) == [0, 1, 3, 6, 2, 4, 5]


def test_small_sizes(O):
    # up to F objects: one leaf root in insertion order; F+1: one split
    for n in (1, 2, 7, 8):
        pts = rand_points(n, 3, seed=n)
        first, count, leaf, level, order = _tree(O, pts, 3, False, 8)
        assert list(leaf) == [1] and list(count) == [n] and list(first) == [0]
        assert list(order) == list(range(n))
    pts = rand_points(9, 3, seed=9)
    first, count, leaf, level, order = _tree(O, pts, 3, False, 8)
    assert list(leaf) == [0, 1, 1] and int(count[0]) == 2
    assert int(count[1]) + int(count[2]) == 9 and min(count[1:]) >= 4
    assert sorted(order) == list(range(9))
    first, count, leaf, level, order = _tree(O, np.zeros((0, 2), np.float32), 2, False, 8)
    assert len(first) == 0


def _check_structure(first, count, leaf, level, order, n, F):
    m = len(first)
    assert sorted(order.tolist()) == list(range(n)), "insertion order is a permutation"
    assert (np.diff(level.astype(int)) >= 0).all(), "BFS order by depth"
    depth = int(level.max())
    assert (level[leaf == 1] == depth).all(), "one leaf level"
    assert (leaf[level < depth] == 0).all()
    for i in range(m):
        c = int(count[i])
        assert 1 <= c <= F
        if i > 0:
            assert c >= F // 2, "minimum fill below the root"
        elif not leaf[0]:
            assert c >= 2
    # leaves tile [0, n) in BFS order; internal children tile [1, m)
    li = np.flatnonzero(leaf == 1)
    assert int(first[li[0]]) == 0
    assert (first[li[1:]] == first[li[:-1]] + count[li[:-1]]).all()
    assert int(first[li[-1]] + count[li[-1]]) == n
    ii = np.flatnonzero(leaf == 0)
    if len(ii):
        assert int(first[ii[0]]) == 1
        assert (first[ii[1:]] == first[ii[:-1]] + count[ii[:-1]]).all()
        assert int(first[ii[-1]] + count[ii[-1]]) == m


@pytest.mark.parametrize("D,is_box,F,n", [(2, False, 128, 30000), (3, False, 16, 20000),
                                          (2, True, 32, 20000), (3, True, 128, 40000),
                                          (2, False, 4, 3000), (3, False, 5, 2000)])
def test_structure_and_search(O, D, is_box, F, n):
    objs = rand_boxes(n, D, seed=n + F) if is_box else rand_points(n, D, seed=n + F,
                                                                    dup_frac=0.2)
    # insertion in Hilbert order, as the build does
    t = O.build_packed(objs, D, is_box, F, O.KEY_B16, layout="guttman")
    _check_structure(t.first, t.count, t.leaf, t.level, t.order, n, F)
    # exact covers: every node box is the min / max of its entries
    E = 2 * D if is_box else D
    lo = t.sobjs[:, :D]
    hi = t.sobjs[:, D:] if is_box else t.sobjs
    for i in range(t.n_nodes):
        f, c = int(t.first[i]), int(t.count[i])
        if t.leaf[i]:
            elo, ehi = lo[f:f + c], hi[f:f + c]
        else:
            elo, ehi = t.boxes[f:f + c, :D], t.boxes[f:f + c, D:]
        assert_bits_equal(t.boxes[i, :D], elo.min(0), "cover lo %d" % i)
        assert_bits_equal(t.boxes[i, D:], ehi.max(0), "cover hi %d" % i)
    assert t.sobjs.shape == (n, E)
    assert_bits_equal(t.sobjs, np.asarray(objs, np.float32)[t.perm], "objects in depth-first order")
    q = rand_queries(300, D, seed=F, objs=objs)
    got = t.search_dfs(q, want_ids=True)
    exp = O.search_brute(t.sobjs, D, is_box, q, want_ids=True)
    assert_bits_equal(got["counts"], exp["counts"], "DFS counts vs brute force")
    assert_bits_equal(got["hashes"], exp["hashes"], "DFS hashes vs brute force")
    assert_bits_equal(got["ids"], exp["ids"], "DFS ids vs brute force")


def test_insertion_order_matters(O):
    # the tree depends on the insertion order (unlike the packed layouts)
    pts = rand_points(5000, 2, seed=5)
    a = _tree(O, pts, 2, False, 16)
    b = _tree(O, pts[::-1].copy(), 2, False, 16)
    assert len(a[0]) != len(b[0]) or not np.array_equal(a[4], b[4][::-1])

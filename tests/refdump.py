"""Writer of the reference's MPHR dump format, for import tests.

Restates the byte layout MPHR::DumpToFile produces (src/tree/mphr.cpp:167-194):
u32 partitions, i64 root offset per partition, u32 node count, then Node_SOA
records (src/node/node_soa.h:40-55): points[2*D*F] f32 (dimension-major, lo
dimensions then hi), index[F] i64, child_offset[F] i64, node_type i32, level i32,
branch_count u32, padded to 8 bytes.  Each partition is a bottom-up tree over a
contiguous rank range (src/tree/mphr.cpp:67-108) with nodes root first; leaf
entries carry their global 1-based rank as index.  Only what the importer reads
(node types, counts, leaf points and indexes) has to be exact; internal nodes
get their children's boxes and last indexes."""
import numpy as np

LEAF, EXTENDLEAF, INTERNAL = 1, 2, 3


def node_dtype(D, F):
    raw = 2 * D * F * 4 + 16 * F + 12
    size = (raw + 7) // 8 * 8
    return np.dtype({"names": ["points", "index", "child_offset", "node_type", "level",
                               "branch_count"],
                     "formats": [("<f4", (2 * D * F,)), ("<i8", (F,)), ("<i8", (F,)), "<i4",
                                 "<i4", "<u4"],
                     "offsets": [0, 8 * D * F, 8 * D * F + 8 * F, 8 * D * F + 16 * F,
                                 8 * D * F + 16 * F + 4, 8 * D * F + 16 * F + 8],
                     "itemsize": size})


def write(path, sobjs, D, F=192, partitions=1, boxes=False, corrupt=None):
    """sobjs: objects in rank order (N x D points or N x 2D boxes)."""
    sobjs = np.ascontiguousarray(sobjs, np.float32)
    n = len(sobjs)
    lo = sobjs[:, :D]
    hi = sobjs[:, D:] if boxes else sobjs[:, :D]
    dt = node_dtype(D, F)
    nodes, roots = [], []
    chunk = n // partitions
    start, end = 0, chunk + n % partitions
    for p in range(partitions):
        # levels bottom-up over ranks [start, end): leaves of F entries
        ent_lo, ent_hi = lo[start:end], hi[start:end]
        ent_idx = np.arange(start + 1, end + 1, dtype=np.int64)
        levels = []
        cur_lo, cur_hi, cur_idx, cnt = ent_lo, ent_hi, ent_idx, len(ent_idx)
        leaf = True
        while True:
            m = -(-cnt // F)
            lvl = []
            for j in range(m):
                a, b = j * F, min(cnt, (j + 1) * F)
                lvl.append((cur_lo[a:b], cur_hi[a:b], cur_idx[a:b], leaf))
            levels.append(lvl)
            if m == 1:
                break
            cur_lo = np.array([x[0].min(0) for x in lvl], np.float32)
            cur_hi = np.array([x[1].max(0) for x in lvl], np.float32)
            cur_idx = np.array([x[2][-1] for x in lvl], np.int64)
            cnt, leaf = m, False
        levels = levels[::-1]  # root first
        roots.append(len(nodes))
        h = len(levels)
        for li, lvl in enumerate(levels):
            for (l_, h_, i_, is_leaf) in lvl:
                rec = np.zeros(1, dt)[0]
                c = len(i_)
                for d in range(D):
                    rec["points"][d * F: d * F + c] = l_[:, d]
                    rec["points"][(D + d) * F: (D + d) * F + c] = h_[:, d]
                rec["index"][:c] = i_
                rec["node_type"] = LEAF if is_leaf else (EXTENDLEAF if li == h - 2 else INTERNAL)
                rec["level"] = li
                rec["branch_count"] = c
                nodes.append(rec)
        start, end = end, end + chunk
    arr = np.array(nodes, dtype=dt)
    if corrupt == "duplicate":
        leaves = np.nonzero(arr["node_type"] == LEAF)[0]
        arr["index"][leaves[-1], 0] = arr["index"][leaves[0], 0]
    if corrupt == "count":
        arr["branch_count"][0] = F + 1
    with open(path, "wb") as f:
        f.write(np.uint32(partitions).tobytes())
        f.write(np.array(roots, np.int64).tobytes())
        f.write(np.uint32(len(arr)).tobytes())
        f.write(arr.tobytes())
        if corrupt == "truncated":
            f.truncate(f.tell() - 7)

"""ctypes front end of libqcp.so (the MI355X Hilbert bulk-load + range-search
library).  Mirrors the reference's tree API (src/tree/tree.h:15-155):

    Index.build(...)      ~ tree::Tree::Build (src/tree/mphr.cpp:35-136, bvh.cpp:25-63)
    Index.search(...)     ~ tree::Tree::Search (src/tree/mphr.cpp:196-272)
    Index.save/.load      ~ DumpToFile / DumpFromFile (src/tree/mphr.cpp:138-194)

Nothing here computes: every call goes through the C ABI of the gfx950 library,
and loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libqcp.so")

LAYOUT = {"mphr": 0, "rtree": 1, "bvh": 2, "guttman": 3}
KEY = {"ref": 0, "b16": 1}
SRC = {"lo": 0, "center": 1}
OBJ = {"points": 0, "boxes": 1}
MODE = {"persistent": 0, "plain": 1, "hybrid": 2}
OUT_COUNTS, OUT_HASH, OUT_IDS, OUT_VISITS = 1, 2, 4, 8

STATUS = {0: "ok", -1: "invalid argument", -2: "out of memory", -3: "HIP runtime error",
          -4: "I/O error", -5: "malformed index file", -6: "buffer too small",
          -7: "coordinate outside the key quantiser's domain", -8: "no GPU"}


class QcpError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s: %s (%d)" % (what, STATUS.get(status, "?"), status))


class BuildCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("dims", "fanout", "layout", "key_mode", "key_src", "obj_kind", "device", "flags",
                 "order", "clusters", "kmeans_iters", "partitions")]

ORDER = {"hilbert": 0, "kmeans": 1}
# qcp_build_cfg.flags
BUILD_GRAPH, BUILD_NO_PERM, BUILD_SORT_INDEX, BUILD_SORT_OBJECTS = 1, 2, 4, 8
SORT = {None: 0, "auto": 0, "index": BUILD_SORT_INDEX, "objects": BUILD_SORT_OBJECTS}


def build_flags(graph=False, sort=None, keep_perm=True):
    """qcp_build_cfg.flags: graph-captured build, the key sort's form (None = the
    library's choice by size, "index" = (key, input index) + gather, "objects" =
    objects carried through the passes) and whether the index keeps the input
    index of each rank."""
    return ((BUILD_GRAPH if graph else 0) | SORT[sort] |
            (0 if keep_perm else BUILD_NO_PERM))


class IndexInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("n_nodes", ctypes.c_uint64), ("height", ctypes.c_uint32),
                ("dims", ctypes.c_uint32), ("fanout", ctypes.c_uint32), ("layout", ctypes.c_uint32),
                ("obj_kind", ctypes.c_uint32), ("key_mode", ctypes.c_uint32),
                ("key_src", ctypes.c_uint32), ("key_bits", ctypes.c_uint32),
                ("level_count", ctypes.c_uint64 * 32), ("tree_bytes", ctypes.c_uint64),
                ("build_ms", ctypes.c_double), ("build_phase_ms", ctypes.c_double * 6),
                ("partitions", ctypes.c_uint32), ("kmeans_iterations", ctypes.c_uint32),
                ("kmeans_ms", ctypes.c_double * 2), ("sort_form", ctypes.c_uint32),
                ("has_perm", ctypes.c_uint32)]


class SearchCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("mode", "outputs", "plain_batch", "block_waves", "grid_blocks", "leaf_chunks",
                 "donate", "order", "max_waiters", "hunger_period", "donate_height",
                 "graph", "contain", "cpu_threads", "handoff_level", "hybrid_batch",
                 "plain_waves")]


class SearchStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("total_hits", ctypes.c_uint64), ("internal_entries", ctypes.c_uint64),
                ("leaf_entries", ctypes.c_uint64), ("donated_items", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("contained_runs", ctypes.c_uint64),
                ("search_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneEntry(ctypes.Structure):
    _fields_ = [("fanout", ctypes.c_uint32), ("block_waves", ctypes.c_uint32),
                ("leaf_chunks", ctypes.c_uint32), ("donate", ctypes.c_uint32),
                ("kernel_ms", ctypes.c_double)]


class HybridTuneEntry(ctypes.Structure):
    _fields_ = [("handoff_level", ctypes.c_uint32), ("hybrid_batch", ctypes.c_uint32),
                ("total_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("launches", ctypes.c_uint64)]


class ServeCfg(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("result_budget", ctypes.c_uint64), ("max_requests", ctypes.c_uint32),
                ("grid_blocks", ctypes.c_uint32), ("contain", ctypes.c_uint32),
                ("outputs", ctypes.c_uint32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("deferred_by_budget", ctypes.c_uint64),
                ("max_batch_seen", ctypes.c_uint64), ("max_inflight_seen", ctypes.c_uint64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("bytes_peak", ctypes.c_uint64), ("ids_served", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
P = ctypes.c_void_p


def lib():
    """Load libqcp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqcp.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.qcp_strerror.restype = ctypes.c_char_p
        L.qcp_strerror.argtypes = [ctypes.c_int]
        L.qcp_device_count.restype = ctypes.c_int
        L.qcp_build.argtypes = [ctypes.POINTER(BuildCfg), P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_build_device.argtypes = [ctypes.POINTER(BuildCfg), P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_build_ranked.argtypes = [ctypes.POINTER(BuildCfg), P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_read_reference_mphr.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint32, P, ctypes.c_uint64,
                                              ctypes.POINTER(ctypes.c_uint64)]
        L.qcp_rebuild_device.argtypes = [P, P]
        L.qcp_index_destroy.argtypes = [P]
        L.qcp_index_destroy.restype = None
        L.qcp_index_info_get.argtypes = [P, ctypes.POINTER(IndexInfo)]
        L.qcp_export.argtypes = [P, P, P, P, P, P, P, P]
        L.qcp_save.argtypes = [P, ctypes.c_char_p]
        L.qcp_load.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(P)]
        L.qcp_batch_create.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_batch_create_device.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_batch_destroy.argtypes = [P]
        L.qcp_batch_destroy.restype = None
        L.qcp_batch_run.argtypes = [P, ctypes.POINTER(SearchCfg), ctypes.POINTER(SearchStats)]
        L.qcp_pipeline_run.argtypes = [P, P, P, ctypes.POINTER(SearchCfg),
                                       ctypes.POINTER(SearchStats)]
        L.qcp_batch_results.argtypes = [P, P, P, P, P, ctypes.c_uint64, P]
        L.qcp_batch_ids.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64, P, P, ctypes.c_uint64]
        L.qcp_batch_total_hits.argtypes = [P]
        L.qcp_batch_total_hits.restype = ctypes.c_uint64
        L.qcp_search.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(SearchCfg), P, P, P, P,
                                 ctypes.c_uint64, ctypes.POINTER(SearchStats)]
        L.qcp_autotune_hybrid.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(SearchCfg),
                                          ctypes.POINTER(SearchCfg),
                                          ctypes.POINTER(HybridTuneEntry), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint32)]
        L.qcp_autotune.argtypes = [ctypes.POINTER(BuildCfg), P, ctypes.c_uint64, P,
                                   ctypes.c_uint64, P, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.POINTER(P), ctypes.POINTER(SearchCfg),
                                   ctypes.POINTER(TuneEntry), ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32)]
        L.qcp_device_alloc.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(P)]
        L.qcp_device_free.argtypes = [P]
        L.qcp_copy_to_device.argtypes = [P, P, ctypes.c_uint64]
        L.qcp_copy_to_host.argtypes = [P, P, ctypes.c_uint64]
        L.qcp_server_start.argtypes = [P, ctypes.POINTER(ServeCfg), ctypes.POINTER(P)]
        L.qcp_server_submit.argtypes = [P, P, ctypes.POINTER(ctypes.c_uint64)]
        L.qcp_server_result.argtypes = [P, ctypes.c_uint64, P, P, P]
        L.qcp_server_wait.argtypes = [P, ctypes.c_uint64, P, P, P]
        L.qcp_server_stats.argtypes = [P, ctypes.POINTER(ServeStats)]
        L.qcp_server_stop.argtypes = [P]
        L.qcp_server_replay.argtypes = [P, P, P, ctypes.c_uint64, P, P, P]
        L.qcp_server_replay_ids.argtypes = [P, P, P, ctypes.c_uint64, P, P, P, P, P,
                                            ctypes.c_uint64]
        L.qcp_server_result_ids.argtypes = [P, ctypes.c_uint64, P, ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.qcp_server_release.argtypes = [P, ctypes.c_uint64]
        L.qcp_search_noalloc.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(SearchCfg), P, P]
        L.qcp_probe_read_gbps.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.qcp_philox4x32_10.argtypes = [P, P, P]
        L.qcp_philox4x32_10.restype = None
        for fn in ("qcp_gen_words", "qcp_gen_uniform", "qcp_gen_normal"):
            getattr(L, fn).argtypes = [P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_uint64]
            getattr(L, fn).restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(st, what):
    if st != 0:
        raise QcpError(st, what)


def device_count():
    return lib().qcp_device_count()


def read_reference_mphr(path, dims=3, leaf_degree=192, obj_kind="points"):
    """Objects of a reference MPHR dump in its rank order (host only)."""
    n = ctypes.c_uint64()
    check(lib().qcp_read_reference_mphr(path.encode(), dims, leaf_degree, OBJ[obj_kind], None, 0,
                                        ctypes.byref(n)), "qcp_read_reference_mphr")
    E = dims * (2 if obj_kind == "boxes" else 1)
    out = np.zeros((max(n.value, 1), E), np.float32)
    check(lib().qcp_read_reference_mphr(path.encode(), dims, leaf_degree, OBJ[obj_kind], _p(out),
                                        n.value, ctypes.byref(n)), "qcp_read_reference_mphr")
    return out[: n.value]


# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A caller-owned HBM buffer holding a copy of a host array."""

    def __init__(self, host, device=0):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.shape, self.dtype = host.shape, host.dtype
        self.ptr = P()
        check(lib().qcp_device_alloc(device, self.nbytes, ctypes.byref(self.ptr)), "qcp_device_alloc")
        check(lib().qcp_copy_to_device(self.ptr, _p(host), self.nbytes), "qcp_copy_to_device")

    def write(self, host):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        check(lib().qcp_copy_to_device(self.ptr, _p(host), self.nbytes), "qcp_copy_to_device")

    def read(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().qcp_copy_to_host(_p(out), self.ptr, self.nbytes), "qcp_copy_to_host")
        return out

    def close(self):
        if self.ptr:
            lib().qcp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """A device-resident tree built by libqcp.so."""

    def __init__(self, handle, D, obj_kind):
        self._h = handle
        self.D = D
        self.obj_kind = obj_kind

    @classmethod
    def build(cls, objs, dims, fanout=64, layout="mphr", key_mode="ref", key_src="lo",
              obj_kind="points", device=0, graph=False, order="hilbert", clusters=0,
              kmeans_iters=0, partitions=1, sort=None, keep_perm=True):
        """Build from host objects (N x D points or N x 2D boxes).  With
        graph=True the device pipeline runs as a captured hipGraph; order="kmeans"
        ranks k-means clusters by their centres' Hilbert keys first; partitions=P
        splits an MPHR into P subtrees over contiguous rank ranges; sort and
        keep_perm as for build_flags."""
        if isinstance(objs, DeviceBuffer):
            return cls.build_device(objs, dims, fanout, layout, key_mode, key_src, obj_kind,
                                    device, graph, partitions, sort, keep_perm)
        objs = np.ascontiguousarray(objs, dtype=np.float32)
        E = dims * (2 if obj_kind == "boxes" else 1)
        n = objs.size // E if objs.size else 0
        cfg = BuildCfg(dims, fanout, LAYOUT[layout], KEY[key_mode], SRC[key_src],
                       OBJ[obj_kind], device, build_flags(graph, sort, keep_perm),
                       ORDER[order], clusters, kmeans_iters, partitions)
        h = P()
        check(lib().qcp_build(ctypes.byref(cfg), _p(objs) if n else None, n, ctypes.byref(h)),
              "qcp_build")
        return cls(h, dims, obj_kind)

    @classmethod
    def build_ranked(cls, objs, dims, fanout=64, layout="mphr", obj_kind="points", device=0,
                     partitions=1):
        """Build from objects already in rank order (entry i gets rank ID i+1)."""
        objs = np.ascontiguousarray(objs, dtype=np.float32)
        E = dims * (2 if obj_kind == "boxes" else 1)
        n = objs.size // E if objs.size else 0
        cfg = BuildCfg(dims, fanout, LAYOUT[layout], 0, 0, OBJ[obj_kind], device, 0, 0, 0, 0,
                       partitions)
        h = P()
        check(lib().qcp_build_ranked(ctypes.byref(cfg), _p(objs) if n else None, n,
                                     ctypes.byref(h)), "qcp_build_ranked")
        return cls(h, dims, obj_kind)

    @classmethod
    def build_device(cls, buf, dims, fanout=64, layout="mphr", key_mode="ref", key_src="lo",
                     obj_kind="points", device=0, graph=False, partitions=1, sort=None,
                     keep_perm=True):
        """Build from objects already resident in HBM (a DeviceBuffer)."""
        E = dims * (2 if obj_kind == "boxes" else 1)
        n = buf.nbytes // (4 * E)
        cfg = BuildCfg(dims, fanout, LAYOUT[layout], KEY[key_mode], SRC[key_src],
                       OBJ[obj_kind], device, build_flags(graph, sort, keep_perm), 0, 0, 0,
                       partitions)
        h = P()
        check(lib().qcp_build_device(ctypes.byref(cfg), buf.ptr, n, ctypes.byref(h)),
              "qcp_build_device")
        return cls(h, dims, obj_kind)

    def rebuild(self, buf):
        """Rebuild in place from a DeviceBuffer of the same size (timed builds)."""
        check(lib().qcp_rebuild_device(self._h, buf.ptr), "qcp_rebuild_device")

    @classmethod
    def load(cls, path, device=0):
        h = P()
        check(lib().qcp_load(path.encode(), device, ctypes.byref(h)), "qcp_load")
        ix = cls(h, 0, "points")
        inf = ix.info()
        ix.D = inf["dims"]
        ix.obj_kind = "boxes" if inf["obj_kind"] == 1 else "points"
        return ix

    def close(self):
        if self._h:
            lib().qcp_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = IndexInfo()
        check(lib().qcp_index_info_get(self._h, ctypes.byref(inf)), "qcp_index_info_get")
        d = {k: getattr(inf, k) for k, _ in IndexInfo._fields_}
        d["level_count"] = [int(x) for x in inf.level_count[: inf.height]]
        d["build_phase_ms"] = list(inf.build_phase_ms)
        d["kmeans_ms"] = list(inf.kmeans_ms)
        return d

    def export(self):
        inf = self.info()
        n, m, D = inf["n"], inf["n_nodes"], inf["dims"]
        E = D * (2 if inf["obj_kind"] == 1 else 1)
        keys = np.zeros(max(n, 1), np.uint64)
        perm = np.zeros(max(n, 1), np.uint32) if inf["has_perm"] else None
        sobjs = np.zeros((max(n, 1), E), np.float32)
        boxes = np.zeros((max(m, 1), 2 * D), np.float32)
        first = np.zeros(max(m, 1), np.uint32)
        count = np.zeros(max(m, 1), np.uint32)
        leaf = np.zeros(max(m, 1), np.uint8)
        check(lib().qcp_export(self._h, _p(keys), _p(perm) if perm is not None else None,
                               _p(sobjs), _p(boxes), _p(first), _p(count), _p(leaf)),
              "qcp_export")
        return {"keys": keys[:n], "perm": perm[:n] if perm is not None else None,
                "sobjs": sobjs[:n], "boxes": boxes[:m],
                "first": first[:m], "count": count[:m], "leaf": leaf[:m], "info": inf}

    def save(self, path):
        check(lib().qcp_save(self._h, path.encode()), "qcp_save")

    def batch(self, queries):
        return Batch(self, queries)

    def search(self, queries, mode="persistent", ids=False, visits=False, plain_batch=128,
               grid_blocks=0, donate=0, chunks=0, hash=True, order=0, block_waves=0,
               contain=False, plain_waves=0, **hybrid):
        with self.batch(queries) as b:
            st = b.run(mode=mode, ids=ids, plain_batch=plain_batch, grid_blocks=grid_blocks,
                       donate=donate, chunks=chunks, hash=hash, order=order,
                       block_waves=block_waves, contain=contain, plain_waves=plain_waves,
                       **hybrid)
            res = b.results(ids=ids, visits=visits)
            res["stats"] = st
            return res


class Batch:
    """A query set resident in HBM, searchable repeatedly (bench timing)."""

    def __init__(self, index, queries):
        self.index = index
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 2 * index.D)
        self.q = q.shape[0]
        self._h = P()
        check(lib().qcp_batch_create(index._h, _p(q) if self.q else None, self.q,
                                     ctypes.byref(self._h)), "qcp_batch_create")

    def close(self):
        if self._h:
            lib().qcp_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, mode="persistent", ids=False, plain_batch=128, grid_blocks=0, donate=0,
            hash=True, chunks=0, order=0, max_waiters=0, hunger_period=0, donate_height=0,
            graph=False, block_waves=0, contain=False, cpu_threads=0, handoff_level=0,
            hybrid_batch=0, plain_waves=0):
        """One search of the resident batch.  donate=1 enables subtree donation
        (work stealing) in the persistent form; chunks = 64-entry leaf chunks in
        flight per wave (0 = 8); order=0 processes the queries in Hilbert
        order of their centres, one range per XCD, order=1 as given;
        block_waves = waves per workgroup (0 = 8); contain=True reports subtrees
        whose box lies inside the query without visiting them (same results);
        plain_waves = 3 runs the plain form with the reference's 192-thread workgroups.
        mode="hybrid": host threads walk the levels above handoff_level (0 = the
        leaves) and the GPU searches the subtrees they hand over."""
        outputs = OUT_COUNTS | (OUT_HASH if hash else 0) | (OUT_IDS if ids else 0)
        cfg = SearchCfg(MODE[mode], outputs, plain_batch, block_waves, grid_blocks, chunks,
                        donate, order, max_waiters, hunger_period, donate_height,
                        1 if graph else 0, 1 if contain else 0, cpu_threads, handoff_level,
                        hybrid_batch, plain_waves)
        st = SearchStats()
        check(lib().qcp_batch_run(self._h, ctypes.byref(cfg), ctypes.byref(st)), "qcp_batch_run")
        self._ids = ids
        self._mode = mode
        return st.as_dict()

    def pipeline(self, buf, block_waves=0, chunks=0, donate=0, order=0, contain=False):
        """Rebuild the index from a DeviceBuffer and search this batch as one
        captured hipGraph (replayed while buf and the configuration stay)."""
        cfg = SearchCfg(MODE["persistent"], OUT_COUNTS | OUT_HASH, 0, block_waves, 0, chunks,
                        donate, order, 0, 0, 0, 1, 1 if contain else 0)
        st = SearchStats()
        check(lib().qcp_pipeline_run(self.index._h, buf.ptr, self._h, ctypes.byref(cfg),
                                     ctypes.byref(st)), "qcp_pipeline_run")
        self._ids = False
        self._mode = "persistent"
        return st.as_dict()

    def offsets_range(self, q0, q1):
        """Offsets (relative to query q0) of the ID lists of queries [q0, q1), without
        copying the IDs themselves."""
        off = np.zeros(q1 - q0 + 1, np.uint64)
        check(lib().qcp_batch_ids(self._h, q0, q1, _p(off), None, 0), "qcp_batch_ids")
        return off

    def ids_range(self, q0, q1):
        """Offsets (relative) and sorted rank-ID lists of queries [q0, q1)."""
        off = np.zeros(q1 - q0 + 1, np.uint64)
        check(lib().qcp_batch_ids(self._h, q0, q1, _p(off), None, 0), "qcp_batch_ids")
        ids = np.zeros(max(int(off[-1]), 1), np.uint32)
        check(lib().qcp_batch_ids(self._h, q0, q1, None, _p(ids), int(off[-1])), "qcp_batch_ids")
        return off, ids[: int(off[-1])]

    def results(self, ids=False, visits=False):
        q = self.q
        counts = np.zeros(max(q, 1), np.uint32)
        hashes = np.zeros(max(q, 1), np.uint64)
        out = {}
        offsets = idsb = vis = None
        cap = 0
        if ids:
            total = lib().qcp_batch_total_hits(self._h)
            offsets = np.zeros(q + 1, np.uint64)
            idsb = np.zeros(max(int(total), 1), np.uint32)
            cap = int(total)
        if visits:
            vis = np.zeros(max(2 * q, 2), np.uint32)
        check(lib().qcp_batch_results(self._h, _p(counts), _p(hashes), _p(offsets), _p(idsb), cap,
                                      _p(vis)), "qcp_batch_results")
        out["counts"] = counts[:q]
        out["hashes"] = hashes[:q]
        if ids:
            out["offsets"] = offsets
            out["ids"] = idsb[: int(offsets[-1])]
        if visits:
            out["visits"] = vis[: 2 * q].reshape(q, 2)
        return out


def autotune(objs, dims, sample_queries, fanouts=(32, 64, 128), layout="mphr", key_mode="ref",
             key_src="lo", obj_kind="points", device=0, hash=True):
    """Run the library's autotuner: returns (Index built at the best fanout,
    dict of Batch.run keyword arguments, list of all measurements)."""
    objs_buf = objs if isinstance(objs, DeviceBuffer) else DeviceBuffer(
        np.ascontiguousarray(objs, np.float32), device)
    qs = np.ascontiguousarray(sample_queries, np.float32).reshape(-1, 2 * dims)
    qbuf = DeviceBuffer(qs, device)
    E = dims * (2 if obj_kind == "boxes" else 1)
    n = objs_buf.nbytes // (4 * E)
    cfg = BuildCfg(dims, 0, LAYOUT[layout], KEY[key_mode], SRC[key_src], OBJ[obj_kind], device, 0)
    fs = np.ascontiguousarray(fanouts, np.uint32)
    h = P()
    best = SearchCfg()
    cap = 4096
    table = (TuneEntry * cap)()
    tlen = ctypes.c_uint32()
    outputs = OUT_COUNTS | (OUT_HASH if hash else 0)
    check(lib().qcp_autotune(ctypes.byref(cfg), objs_buf.ptr, n, qbuf.ptr, len(qs), _p(fs),
                             len(fs), outputs, ctypes.byref(h), ctypes.byref(best), table, cap,
                             ctypes.byref(tlen)), "qcp_autotune")
    qbuf.close()
    if objs_buf is not objs:
        objs_buf.close()
    ix = Index(h, dims, obj_kind)
    kw = {"block_waves": best.block_waves, "chunks": best.leaf_chunks, "donate": best.donate}
    rows = [{"fanout": e.fanout, "block_waves": e.block_waves, "leaf_chunks": e.leaf_chunks,
             "donate": e.donate, "kernel_ms": e.kernel_ms} for e in table[: tlen.value]]
    return ix, kw, rows


def autotune_hybrid(index, sample_queries, cpu_threads=0, hash=True, block_waves=0, chunks=0,
                    donate=0, contain=False, device=0):
    """Run the library's co-processing tuner on `index` (on `device`): returns
    (dict of Batch.run keyword arguments for the fastest hybrid split, list of
    all measurements)."""
    qs = np.ascontiguousarray(sample_queries, np.float32).reshape(-1, 2 * index.D)
    qbuf = DeviceBuffer(qs, device)
    base = SearchCfg()
    base.mode = 2
    base.outputs = OUT_COUNTS | (OUT_HASH if hash else 0)
    base.block_waves = block_waves
    base.leaf_chunks = chunks
    base.donate = donate
    base.contain = 1 if contain else 0
    base.cpu_threads = cpu_threads
    best = SearchCfg()
    cap = 256
    table = (HybridTuneEntry * cap)()
    tlen = ctypes.c_uint32()
    try:
        check(lib().qcp_autotune_hybrid(index._h, qbuf.ptr, len(qs), ctypes.byref(base),
                                        ctypes.byref(best), table, cap, ctypes.byref(tlen)),
              "qcp_autotune_hybrid")
    finally:
        qbuf.close()
    kw = {"mode": "hybrid", "handoff_level": best.handoff_level, "hybrid_batch": best.hybrid_batch,
          "cpu_threads": cpu_threads, "block_waves": block_waves, "chunks": chunks,
          "donate": donate, "contain": contain}
    rows = [{"handoff_level": e.handoff_level, "hybrid_batch": e.hybrid_batch,
             "total_ms": e.total_ms, "device_ms": e.device_ms, "launches": e.launches}
            for e in table[: tlen.value]]
    return kw, rows


class Server:
    """Continuous-batching server over a resident index: counts + ID-set hash,
    or with ids=True also the sorted rank-ID lists (admission then charged the
    lists' bytes, sized by each batch's count pass)."""

    def __init__(self, index, slots=4, max_batch=4096, result_budget=0, max_requests=0,
                 grid_blocks=0, contain=False, ids=False):
        self.index = index
        self.ids = ids
        outputs = OUT_COUNTS | OUT_HASH | (OUT_IDS if ids else 0)
        cfg = ServeCfg(slots, max_batch, result_budget, max_requests, grid_blocks,
                       1 if contain else 0, outputs)
        self._h = P()
        check(lib().qcp_server_start(index._h, ctypes.byref(cfg), ctypes.byref(self._h)),
              "qcp_server_start")

    def submit(self, box):
        b = np.ascontiguousarray(box, np.float32)
        t = ctypes.c_uint64()
        check(lib().qcp_server_submit(self._h, _p(b), ctypes.byref(t)), "qcp_server_submit")
        return t.value

    def wait(self, ticket):
        """(count, hash, latency_us); with ID-list service the IDs are still to be
        read with result_ids (or the ticket released)."""
        c = ctypes.c_uint32()
        h = ctypes.c_uint64()
        lat = ctypes.c_double()
        check(lib().qcp_server_wait(self._h, ticket, ctypes.byref(c), ctypes.byref(h),
                                    ctypes.byref(lat)), "qcp_server_wait")
        return c.value, h.value, lat.value

    def result_ids(self, ticket):
        """The sorted rank-ID list of a finished request (frees its entry)."""
        n = ctypes.c_uint64()
        check(lib().qcp_server_result_ids(self._h, ticket, None, 0, ctypes.byref(n)),
              "qcp_server_result_ids")
        out = np.zeros(max(n.value, 1), np.uint32)
        check(lib().qcp_server_result_ids(self._h, ticket, _p(out), n.value, ctypes.byref(n)),
              "qcp_server_result_ids")
        return out[: n.value]

    def release(self, ticket):
        check(lib().qcp_server_release(self._h, ticket), "qcp_server_release")

    def replay(self, boxes, arrival_s, ids_cap=None):
        """Submit boxes[i] at arrival_s[i] (native threads), read every result out;
        returns (counts, hashes, latency_us), plus (offsets, ids) with ID-list
        service (ids_cap = 0: lists served but not copied)."""
        boxes = np.ascontiguousarray(boxes, np.float32)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        n = len(arr)
        counts = np.zeros(max(n, 1), np.uint32)
        hashes = np.zeros(max(n, 1), np.uint64)
        lat = np.zeros(max(n, 1), np.float64)
        if not self.ids:
            check(lib().qcp_server_replay(self._h, _p(boxes), _p(arr), n, _p(counts),
                                          _p(hashes), _p(lat)), "qcp_server_replay")
            return counts[:n], hashes[:n], lat[:n]
        offsets = np.zeros(n + 1, np.uint64)
        cap = 0 if ids_cap is None else int(ids_cap)
        idsb = np.zeros(max(cap, 1), np.uint32) if cap else None
        check(lib().qcp_server_replay_ids(self._h, _p(boxes), _p(arr), n, _p(counts),
                                          _p(hashes), _p(lat), _p(offsets), _p(idsb), cap),
              "qcp_server_replay_ids")
        return (counts[:n], hashes[:n], lat[:n], offsets,
                idsb[: int(offsets[-1])] if idsb is not None else None)

    def stats(self):
        st = ServeStats()
        check(lib().qcp_server_stats(self._h, ctypes.byref(st)), "qcp_server_stats")
        return st.as_dict()

    def stop(self):
        if self._h:
            h, self._h = self._h, None
            check(lib().qcp_server_stop(h), "qcp_server_stop")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


NOALLOC_MAX_QUERIES = 262144


def search_noalloc(index, d_queries, q, d_counts, d_hashes=None, hash=True, donate=0, chunks=0,
                   block_waves=0, order=0, contain=False, grid_blocks=0):
    """The persistent search through the entry point that allocates nothing:
    d_queries / d_counts / d_hashes are DeviceBuffers the caller made (queries
    q x 2D floats, counts u32[q], hashes u64[q] when hash)."""
    outputs = OUT_COUNTS | (OUT_HASH if hash else 0)
    cfg = SearchCfg(MODE["persistent"], outputs, 0, block_waves, grid_blocks, chunks, donate,
                    order, 0, 0, 0, 0, 1 if contain else 0, 0, 0, 0, 0)
    check(lib().qcp_search_noalloc(index._h, d_queries.ptr if q else None, q, ctypes.byref(cfg),
                                   d_counts.ptr if q else None,
                                   d_hashes.ptr if d_hashes is not None else None),
          "qcp_search_noalloc")


def probe_read_gbps(window_bytes, rounds, device=0):
    """(GB/s, ms) of dword-load streaming over a window of window_bytes (the
    memory level that holds it: L2, Infinity Cache or HBM)."""
    g, m = ctypes.c_double(), ctypes.c_double()
    check(lib().qcp_probe_read_gbps(device, window_bytes, rounds, ctypes.byref(g),
                                    ctypes.byref(m)), "qcp_probe_read_gbps")
    return g.value, m.value


def poisson_arrivals(n, qps, seed=1234, stream=3):
    """Arrival times (s) of a Poisson process of rate qps: cumulative
    exponential gaps -ln(1-u)/qps from Philox stream `stream`."""
    u = gen_uniform(n, seed, stream).astype(np.float64)
    return np.cumsum(-np.log1p(-u) / qps)


# ---------------------------------------------------------------------------
# seeded generator (Philox4x32-10) and the benchmark workloads
def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().qcp_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def gen_words(n, seed, stream, offset=0):
    out = np.empty(max(n, 1), np.uint32)
    lib().qcp_gen_words(_p(out), n, seed, stream, offset)
    return out[:n]


def gen_uniform(n, seed, stream, offset=0):
    out = np.empty(max(n, 1), np.float32)
    lib().qcp_gen_uniform(_p(out), n, seed, stream, offset)
    return out[:n]


def gen_normal(n, seed, stream, offset=0):
    out = np.empty(max(n, 1), np.float64)
    lib().qcp_gen_normal(_p(out), n, seed, stream, offset)
    return out[:n]


SEED = 1234        # echoes the reference generator's seed (generator/cuRAND.cpp:49)
CLIP_HI = np.float32(1.0) - np.float32(2.0 ** -24)
# streams: 0 data, 1 query centres, 2 query sizes, 3 arrivals, 4 mixture centres,
# 5 mixture components, 6 mixture noise, 7 query components, 8 query noise


def boxes_from(centers, sides):
    """Query boxes lo = c - 0.5f*side, hi = c + 0.5f*side in FP32, not clipped."""
    c = np.asarray(centers, np.float32)
    h = np.float32(0.5) * np.asarray(sides, np.float32)
    return np.ascontiguousarray(np.concatenate([c - h, c + h], axis=1), np.float32)


def uniform_points(n, D, seed=SEED, stream=0):
    return gen_uniform(n * D, seed, stream).reshape(n, D)


def mixture_points(n, D, k, sigma, seed=SEED, comp_stream=5, noise_stream=6, centers=None):
    """n points from a k-component equal-weight Gaussian mixture with the given
    centres (uniform in [0,1)^D from stream 4 when None), clipped to
    [0, 1-2^-24]."""
    if centers is None:
        centers = gen_uniform(k * D, seed, 4).reshape(k, D)
    comp = (gen_words(n, seed, comp_stream).astype(np.uint64) * np.uint64(k)) >> np.uint64(32)
    z = gen_normal(n * D, seed, noise_stream).reshape(n, D)
    p = (centers[comp].astype(np.float64) + sigma * z).astype(np.float32)
    np.clip(p, np.float32(0.0), CLIP_HI, out=p)
    return p, centers


def workload(name, scale=1.0, seed=SEED, nq=None):
    """Synthetic inputs of one benchmark configuration.  `scale` shrinks N and Q
    proportionally (parity runs); 1.0 is the full size; `nq` overrides the
    number of queries."""
    if name == "c1":
        n, q = int(1048576 * scale), max(1, int(1000 * min(1.0, scale * 4)))
        pts = uniform_points(n, 2, seed)
        cen = gen_uniform(q * 2, seed, 1).reshape(q, 2)
        side = np.full((q, 2), np.float32(np.sqrt(np.float32(1e-4))), np.float32)
        return {"objs": pts, "queries": boxes_from(cen, side), "dims": 2, "obj_kind": "points",
                "fanout": 64, "key_mode": "b16", "layout": "mphr"}
    if name == "c2":
        n, q = int(16777216 * scale), max(1, int(100000 * scale))
        pts, centers = mixture_points(n, 3, 64, 0.01, seed)
        qc, _ = mixture_points(q, 3, 64, 0.01, seed, comp_stream=7, noise_stream=8,
                               centers=centers)
        u = gen_uniform(q, seed, 2).astype(np.float64)
        sel = 10.0 ** (-6.0 + 4.0 * u)
        side = np.repeat((sel ** (1.0 / 3.0)).astype(np.float32)[:, None], 3, axis=1)
        return {"objs": pts, "queries": boxes_from(qc, side), "dims": 3, "obj_kind": "points",
                "fanout": 64, "key_mode": "ref", "layout": "mphr"}
    if name == "c3":
        n, q = int(67108864 * scale), max(1, int(1000000 * scale))
        lo = gen_uniform(n * 2, seed, 0).reshape(n, 2)
        u = gen_uniform(n * 2, seed, 9).astype(np.float64)
        side = (-1e-4 * np.log1p(-u)).astype(np.float32).reshape(n, 2)
        objs = np.concatenate([lo, lo + side], axis=1).astype(np.float32)
        cen = gen_uniform(q * 2, seed, 1).reshape(q, 2)
        heavy = gen_uniform(q, seed, 2) < np.float32(0.1)
        s = np.where(heavy, np.float32(np.sqrt(1e-3)), np.float32(1e-3)).astype(np.float32)
        return {"objs": objs, "queries": boxes_from(cen, np.repeat(s[:, None], 2, 1)), "dims": 2,
                "obj_kind": "boxes", "fanout": 128, "key_mode": "ref", "key_src": "center",
                "layout": "rtree"}
    if name == "c4":
        n, q = int(33554432 * scale), max(1, int(100000 * scale))
        q = nq or q
        pts = uniform_points(n, 3, seed)
        cen = gen_uniform(q * 3, seed, 1).reshape(q, 3)
        u = gen_uniform(q, seed, 2).astype(np.float64)
        sel = 10.0 ** (-7.0 + 4.0 * u)
        side = np.repeat((sel ** (1.0 / 3.0)).astype(np.float32)[:, None], 3, axis=1)
        return {"objs": pts, "queries": boxes_from(cen, side), "dims": 3, "obj_kind": "points",
                "fanout": 64, "key_mode": "ref", "layout": "mphr"}
    if name == "c5":
        n, q = int(134217728 * scale), max(1, int(500000 * scale))
        nh = int(n * 0.8)
        hot_c = gen_uniform(16 * 2, seed, 4).reshape(16, 2)
        w = np.arange(1, 17, dtype=np.float64) ** -1.2
        cdf = np.cumsum(w / w.sum())
        u = gen_uniform(nh, seed, 5).astype(np.float64)
        comp = np.minimum(np.searchsorted(cdf, u, side="right"), 15)
        z = gen_normal(nh * 2, seed, 6).reshape(nh, 2)
        hot = (hot_c[comp].astype(np.float64) + 0.002 * z).astype(np.float32)
        cold = gen_uniform((n - nh) * 2, seed, 0).reshape(n - nh, 2)
        pts = np.concatenate([hot, cold]).astype(np.float32)
        np.clip(pts, np.float32(0.0), CLIP_HI, out=pts)
        pick = (gen_words(q, seed, 7).astype(np.uint64) * np.uint64(n)) >> np.uint64(32)
        cen = pts[pick.astype(np.int64)]
        sides = (np.float32(1e-4) + gen_uniform(q * 2, seed, 2).reshape(q, 2) *
                 np.float32(1e-2 - 1e-4)).astype(np.float32)
        return {"objs": pts, "queries": boxes_from(cen, sides), "dims": 2, "obj_kind": "points",
                "fanout": 64, "key_mode": "ref", "layout": "mphr"}
    raise ValueError(name)

"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared symbol, error paths return status codes, the Philox generator matches
its published known-answer vectors, and the workload definitions."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "qcp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qcp_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(qcp):
    L = qcp.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings(qcp):
    L = qcp.lib()
    for s in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert L.qcp_strerror(s)
    assert L.qcp_strerror(0) == b"ok"


def test_invalid_arguments_return_status(qcp):
    import ctypes

    L = qcp.lib()
    h = ctypes.c_void_p()
    assert L.qcp_build(None, None, 0, ctypes.byref(h)) != 0
    cfg = qcp.BuildCfg(5, 64, 0, 0, 0, 0, 0, 0)   # dims 5 is rejected
    assert L.qcp_build(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    cfg = qcp.BuildCfg(2, 1, 0, 0, 0, 0, 0, 0)    # fanout 1 is rejected
    assert L.qcp_build(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    cfg = qcp.BuildCfg(2, 64, 0, 0, 1, 0, 0, 0)   # centre key source needs boxes
    assert L.qcp_build(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    assert L.qcp_batch_run(None, None, None) == -1
    assert L.qcp_load(b"/nonexistent/file.qcp", 0, ctypes.byref(h)) == -4


def test_load_rejects_garbage(qcp, tmp_path):
    import ctypes

    p = tmp_path / "junk.qcp"
    p.write_bytes(b"NOTANIDX" + b"\x00" * 64)
    h = ctypes.c_void_p()
    assert qcp.lib().qcp_load(str(p).encode(), 0, ctypes.byref(h)) == -5
    p.write_bytes(b"QCPIDX01")   # truncated header
    assert qcp.lib().qcp_load(str(p).encode(), 0, ctypes.byref(h)) == -5


def test_device_count_without_gpu(qcp):
    assert qcp.device_count() >= 0


def test_philox_known_answers(qcp):
    """Random123 known-answer vectors for Philox4x32-10."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, exp in kat:
        assert list(qcp.philox(ctr, key)) == exp


def test_generator_stream_layout(qcp):
    w = qcp.gen_words(1000, 1234, 3)
    # word i = output word i%4 of block i/4
    for i in (0, 1, 5, 998):
        blk = i // 4
        o = qcp.philox([blk & 0xFFFFFFFF, blk >> 32, 0, 0], [1234, 3])
        assert w[i] == o[i % 4]
    # offsets and slices are consistent (order-independent generation)
    assert np.array_equal(qcp.gen_words(100, 1234, 3, offset=7), w[7:107])
    u = qcp.gen_uniform(200000, 99, 0)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    assert np.array_equal(u[:10], ((qcp.gen_words(10, 99, 0) >> 8).astype(np.float32) *
                                   np.float32(2.0 ** -24)))
    z = qcp.gen_normal(200000, 5, 1)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_c1_workload_matches_golden(qcp):
    g = np.load(os.path.join(GOLD, "c1_expected.npz"))
    w = qcp.workload("c1")
    assert w["objs"].shape == (1048576, 2) and w["queries"].shape == (1000, 4)
    assert np.array_equal(w["objs"][:16], g["data_head"])
    assert np.array_equal(w["queries"][:16], g["query_head"])


def test_c1_oracle_matches_golden(qcp, O):
    g = np.load(os.path.join(GOLD, "c1_expected.npz"))
    w = qcp.workload("c1")
    t = O.build_packed(w["objs"], 2, False, 64, O.KEY_B16)
    assert np.array_equal(t.keys[:4096], g["keys_head"])
    assert np.array_equal(t.perm[:4096], g["perm_head"])
    r = t.search_dfs(w["queries"], want_ids=True)
    assert np.array_equal(r["counts"], g["counts"])
    assert np.array_equal(r["hashes"], g["hashes"])
    k = len(g["offsets_head"]) - 1
    assert np.array_equal(r["ids"][: int(r["offsets"][k])], g["ids_head"])
    # expected hits per query ~ N * (s - s^2/4)^2 for side s=0.01 (border effects aside)
    assert 90 < r["counts"].mean() < 115


def test_workload_shapes_small(qcp):
    for name, D, E in (("c2", 3, 3), ("c3", 2, 4), ("c4", 3, 3), ("c5", 2, 2)):
        w = qcp.workload(name, scale=1.0 / 4096)
        assert w["objs"].shape[1] == E and w["queries"].shape[1] == 2 * D
        assert np.isfinite(w["objs"]).all() and np.isfinite(w["queries"]).all()
        if w["obj_kind"] == "points":
            assert w["objs"].min() >= 0.0 and w["objs"].max() < 1.0


def test_cli_generators_match_library(qcp, tmp_path):
    """qcp_cli --gen-data / --gen-queries write the library's Philox streams
    (data stream 0; query centres stream 1, boxes of side selectivity^(1/D))."""
    import subprocess
    cli = os.path.join(os.path.dirname(qcp.LIB_PATH), "qcp_cli")
    data, queries = str(tmp_path / "d.bin"), str(tmp_path / "q.bin")
    subprocess.run([cli, "--gen-data", data, "-d", "1000", "--dims", "2", "--seed", "5"], check=True)
    subprocess.run([cli, "--gen-queries", queries, "-q", "10", "-s", "0.01", "--dims", "2",
                    "--seed", "5"], check=True)
    assert np.array_equal(np.fromfile(data, np.float32), qcp.gen_uniform(2000, 5, 0))
    q = np.fromfile(queries, np.float32).reshape(10, 4)
    cen = qcp.gen_uniform(20, 5, 1).reshape(10, 2)
    h = np.float32(0.5) * np.float32(0.1)
    assert np.array_equal(q[:, :2], cen - h) and np.array_equal(q[:, 2:], cen + h)


@pytest.mark.parametrize("D,parts,boxes", [(3, 1, False), (2, 3, False), (2, 2, True)])
def test_read_reference_mphr_dump(qcp, tmp_path, D, parts, boxes):
    """Objects come back in the dump's rank order from every partition."""
    import refdump
    from helpers import rand_boxes, rand_points
    objs = rand_boxes(5000, D, 3, mean_side=0.01) if boxes else rand_points(5000, D, 3)
    p = str(tmp_path / "mphr.idx")
    refdump.write(p, objs, D, partitions=parts, boxes=boxes)
    got = qcp.read_reference_mphr(p, D, 192, "boxes" if boxes else "points")
    assert got.shape == objs.shape
    assert np.array_equal(got.view(np.uint32), objs.view(np.uint32))


@pytest.mark.parametrize("bad", ["duplicate", "count", "truncated"])
def test_read_reference_mphr_rejects_malformed(qcp, tmp_path, bad):
    import refdump
    from helpers import rand_points
    p = str(tmp_path / "bad.idx")
    refdump.write(p, rand_points(1000, 3, 4), 3, corrupt=bad)
    with pytest.raises(qcp.QcpError) as e:
        qcp.read_reference_mphr(p, 3, 192)
    assert e.value.status == -5


def test_build_flag_rules(qcp):
    """Sort-form and permutation flags are checked before any device work: both
    sort forms at once, the object form for Guttman layouts or the k-means order
    (whose host steps read the sorted input indices), and unknown bits are refused."""
    import ctypes

    L = qcp.lib()
    h = ctypes.c_void_p()
    both = qcp.BUILD_SORT_INDEX | qcp.BUILD_SORT_OBJECTS
    for cfg in (qcp.BuildCfg(2, 64, 0, 0, 0, 0, 0, both),
                qcp.BuildCfg(2, 64, 3, 0, 0, 0, 0, qcp.BUILD_SORT_OBJECTS),
                qcp.BuildCfg(2, 64, 0, 0, 0, 0, 0, qcp.BUILD_SORT_OBJECTS, 1),
                qcp.BuildCfg(2, 64, 0, 0, 0, 0, 0, 16)):
        assert L.qcp_build(ctypes.byref(cfg), None, 0, ctypes.byref(h)) == -1
    assert qcp.build_flags(graph=True, sort="objects", keep_perm=False) == 1 | 8 | 2
    assert qcp.build_flags() == 0

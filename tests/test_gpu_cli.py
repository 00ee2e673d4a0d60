"""The command-line driver (reference-flag compatible): generated data and
queries, every index type (-i r is the reference's Guttman R-tree), the plain /
persistent / co-processing forms, index save + load, and the autotuner, all
checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "query-co-processing-on-heterogeneous-architectures_amd", "_build", "qcp_cli")


def run(*args):
    print("cli", " ".join(args), flush=True)
    r = subprocess.run([CLI] + list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_out(path):
    a = np.loadtxt(path, dtype=np.uint64, ndmin=2)
    return a[:, 0].astype(np.uint32), a[:, 1]


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    data, queries = str(d / "data.bin"), str(d / "queries.bin")
    run("--gen-data", data, "-d", "120000", "--dims", "3", "--seed", "7")
    run("--gen-queries", queries, "-q", "1500", "-s", "1e-3", "--dims", "3", "--seed", "7")
    pts = np.fromfile(data, np.float32).reshape(-1, 3)
    q = np.fromfile(queries, np.float32).reshape(-1, 6)
    return d, data, queries, pts, q


@pytest.mark.parametrize("itype,layout,extra", [
    ("m", "packed", []), ("m", "packed", ["--mode", "plain", "-b", "64"]),
    ("r", "guttman", []), ("pr", "packed", []), ("b", "bvh", []),
    ("h", "packed", ["-p", "4", "-c", "400"]), ("hb", "bvh", ["-l", "2"]),
    ("hr", "guttman", ["-l", "1"]), ("rls", "guttman-input", ["-p", "4"]),
    ("m", "packed", ["--contain"]), ("m", "packed", ["-t", "9"]),
    ("m", "packed", ["-t", "9", "--mode", "plain"])])
def test_cli_search_matches_oracle(gpu, O, files, itype, layout, extra):
    d, data, queries, pts, q = files
    out = str(d / ("out_%s_%s.txt" % (itype, "_".join(x.strip("-") for x in extra))))
    log = run("-i", itype, "--data", data, "--queries", queries, "--dims", "3", "--fanout", "64",
              "--out", out, "-r", "2", *extra)
    assert "Hit :" in log
    if layout == "guttman-input":  # RTree_LS: Guttman insertion in the input order
        t = O.Tree(pts, 3, False, 64, layout="guttman")
    else:
        t = O.build_packed(pts, 3, False, 64, O.KEY_REF, layout=layout)
    exp = t.search_dfs(q)
    c, h = read_out(out)
    assert np.array_equal(c, exp["counts"]) and np.array_equal(h, exp["hashes"])


def test_cli_partitions_need_mphr(gpu, files):
    """-t with an index type other than MPHR is an error, not ignored."""
    d, data, queries, pts, q = files
    r = subprocess.run([CLI, "-i", "b", "-t", "4", "--data", data, "--queries", queries,
                        "--dims", "3"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "invalid argument" in (r.stdout + r.stderr)


def test_cli_index_file_and_autotune(gpu, O, files):
    d, data, queries, pts, q = files
    ixf = str(d / "ix.qcp")
    out1, out2 = str(d / "o1.txt"), str(d / "o2.txt")
    log = run("-i", "m", "--data", data, "--queries", queries, "--dims", "3", "-e", "1",
              "--index", ixf, "--out", out1, "--json")
    assert "autotune:" in log and os.path.exists(ixf)
    # second run loads the saved index (no --data needed)
    run("-i", "m", "--queries", queries, "--dims", "3", "--index", ixf, "--out", out2)
    exp = O.build_packed(pts, 3, False, 64, O.KEY_REF).search_dfs(q)
    for o in (out1, out2):
        c, h = read_out(o)
        assert np.array_equal(c, exp["counts"]) and np.array_equal(h, exp["hashes"])


def test_cli_hybrid_autotune(gpu, O, files):
    """-i h -e 1: the autotuner, then the co-processing tuner picks the hand-off
    level and pipeline block (no -l / -c given); results equal the oracle's."""
    d, data, queries, pts, q = files
    out = str(d / "oh.txt")
    log = run("-i", "h", "--data", data, "--queries", queries, "--dims", "3", "-e", "1",
              "-p", "4", "--out", out)
    assert "autotune:" in log and "hybrid autotune: hand-off level" in log, log
    exp = O.build_packed(pts, 3, False, 64, O.KEY_REF).search_dfs(q)
    c, h = read_out(out)
    assert np.array_equal(c, exp["counts"]) and np.array_equal(h, exp["hashes"])


def test_cli_reference_dump(gpu, O, files):
    """--ref-dump indexes a reference MPHR dump's objects in its rank order."""
    import refdump
    d, data, queries, pts, q = files
    dump, out = str(d / "ref.idx"), str(d / "oref.txt")
    refdump.write(dump, pts, 3, F=192, partitions=3)
    run("-i", "m", "--ref-dump", dump, "--queries", queries, "--dims", "3", "--fanout", "192",
        "--out", out)
    exp = O.Tree(pts, 3, False, 192).search_dfs(q)
    c, h = read_out(out)
    assert np.array_equal(c, exp["counts"]) and np.array_equal(h, exp["hashes"])


@pytest.mark.parametrize("u", ["k", "o"])
def test_cli_cluster_types(gpu, O, files, u):
    """-u k (k-means + Hilbert order) and -u o (input order) give the same
    search results as any order (counts; rank IDs differ by construction)."""
    d, data, queries, pts, q = files
    out = str(d / ("ou_%s.txt" % u))
    run("-i", "m", "-u", u, "--clusters", "16", "--data", data, "--queries", queries, "--dims", "3",
        "--out", out)
    exp = O.build_packed(pts, 3, False, 64, O.KEY_REF).search_dfs(q)
    c, _ = read_out(out)
    assert np.array_equal(c, exp["counts"])


def test_cli_reference_degrees(gpu, O, files):
    """Without --fanout each index type takes the reference's degree: 192 for
    MPHR, 128 for the R-tree (whose depth-first IDs, and so hashes, depend on it)."""
    d, data, queries, pts, q = files
    log = run("-i", "m", "--data", data, "--queries", queries, "--dims", "3")
    assert "fanout 192" in log
    out = str(d / "o_r128.txt")
    log = run("-i", "r", "--data", data, "--queries", queries, "--dims", "3", "--out", out)
    assert "fanout 128" in log
    exp = O.build_packed(pts, 3, False, 128, O.KEY_REF, layout="guttman").search_dfs(q)
    c, h = read_out(out)
    assert np.array_equal(c, exp["counts"]) and np.array_equal(h, exp["hashes"])

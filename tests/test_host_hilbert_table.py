"""The table-driven Hilbert key that the GPU key kernel runs
(csrc/hilbert_key.h, hilbert_key_table: L levels per lookup from a transition
table over the walk's state) against the oracle's restatement of
HilbertMapper::MappingIntoSingle (src/mapper/hilbert_mapper.cpp:15-68) on the
CPU: a compiled program (tests/cpp/test_hilbert_table.cpp) compares them on
every point of small grids and on random and edge coordinates at every bit
width (2-D: 2..31 bits, 3-D: 2..21 bits)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
PROG = os.path.join(CPP, "_build", "test_hilbert_table")


def test_table_key_matches_oracle():
    if not os.path.exists(PROG):
        import oracle_lib

        oracle_lib.lib()  # the oracle library the program links
        subprocess.check_call(["make", "-s", "-C", CPP])
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
    assert int(r.stdout.split()[1]) > 1000000

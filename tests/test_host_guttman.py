"""The library's host Guttman insertion (csrc/guttman.cpp) against the oracle's
restatement on the CPU: a compiled program (tests/cpp/test_guttman_host.cpp)
compares their BFS topologies and depth-first orders over 120 seeded random
cases (2-D / 3-D, points / boxes, capacities 4 to 256, tie-heavy inputs) and
checks that the stored-topology checker refuses corrupted trees."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
PROG = os.path.join(CPP, "_build", "test_guttman_host")


def test_library_guttman_matches_oracle():
    if not os.path.exists(PROG):
        import oracle_lib

        oracle_lib.lib()  # the oracle library the program links
        subprocess.check_call(["make", "-s", "-C", CPP])
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "120 cases, 0 failures" in r.stdout

"""The NBody::KDTree drop-in's host logic on the CPU: the driver program
(tests/cpp/shim_drive.cpp) linked with the drop-in (kdtree_shim.cpp) over a
host stub of the C ABI that the oracle answers, built with the address and
undefined-behaviour sanitizers (tests/cpp/Makefile).  It covers the ID
overwrite and in-place reorder (KDTree.cxx:1291), per-particle calls from
OpenMP threads (test_kdtree.cxx:278-301, :324-341), the target exclusion of
SearchBallPosTagged(tt) (KDLeafNode.cxx:314-333), FOF's linked lists and
check function (KDFOF.cxx:29-153), and ~KDTree restoring the input order
(KDTree.cxx:1340-1356)."""
import os
import subprocess

import pytest

import shimcheck

EXE = os.path.join(shimcheck.CPP, "test_kdtree_shim_cpu")


@pytest.mark.parametrize("periodic", [0, 1])
def test_shim_host_logic_under_sanitizers(tmp_path, periodic):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_kdtree_shim_cpu is not built (run __graft_entry__.build())")
    n, k, b, rball = 3000, 16, 0.06, 0.08
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    pos, _, m = shimcheck.write_input(fin, n, 41 + periodic)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([EXE] + shimcheck.args(fin, fout, k, b, periodic, rball), capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    shimcheck.check(shimcheck.read_output(fout), pos, m, k, b, periodic, rball)


def test_reference_call_lines_compile():
    """tests/cpp/compile_calls.cpp (the reference driver's own call forms and
    the other KDTree.h overloads) compiled with -Wall -Werror by the build."""
    assert os.path.exists(os.path.join(shimcheck.CPP, "compile_calls.o"))

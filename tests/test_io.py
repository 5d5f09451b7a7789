"""parts.out, the reference driver's raw particle file (src/tests/test_kdtree.cxx:60-73, :181-208)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from nbodylib_amd import datagen
from nbodylib_amd.io import read_parts_out, write_parts_out


def test_parts_out_layout_and_round_trip(tmp_path):
    n = 1000
    pos, vel, _ = datagen.clustered(n, 10.0, nhalo=4, seed=51, with_vel=True)
    f = tmp_path / "parts.out"
    write_parts_out(f, pos, vel)
    raw = np.fromfile(f, dtype=np.float64)
    assert raw.size == 6 * n  # no header, six doubles a record
    assert np.array_equal(raw[:6], np.concatenate([pos[0], vel[0]]))
    p2, v2, m2 = read_parts_out(f)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel) and np.all(m2 == 1.0)
    p3, _, _ = read_parts_out(f, n=10)
    assert np.array_equal(p3, pos[:10])
    with pytest.raises(ValueError):
        read_parts_out(f, n=n + 1)
    (tmp_path / "bad.out").write_bytes(b"\0" * 50)
    with pytest.raises(ValueError):
        read_parts_out(tmp_path / "bad.out")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_parts_out_cpp_helpers(tmp_path):
    """NBody::ReadPartsOut / WritePartsOut (include/nbx_kdtree.h) read what
    the Python writer wrote and write the same bytes back."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "parts_io"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "parts_io.cpp"), "-o", str(exe)], check=True)
    pos = datagen.uniform(500, 1.0, 52)
    vel = np.random.default_rng(52).normal(size=(500, 3))
    fin, fout = tmp_path / "in.out", tmp_path / "out.out"
    write_parts_out(fin, pos, vel)
    r = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "500"
    assert fin.read_bytes() == fout.read_bytes()

"""Runs the NBody::KDTree drop-in driver (tests/cpp/shim_drive.cpp) and
checks what it wrote against the CPU oracle.  Shared by the CPU test (the
drop-in over a host stub of the C ABI, under sanitizers) and the GPU test
(the drop-in over libnbx.so, in-process)."""
import os

import numpy as np

import oracle
from nbodylib_amd import datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def write_input(path, n, seed):
    pos = datagen.uniform(n, 1.0, seed)
    vel = np.random.default_rng(seed).normal(size=(n, 3))
    m = datagen.lognormal_masses(n, seed)
    with open(path, "wb") as f:
        np.array([n], dtype=np.int32).tofile(f)
        np.concatenate([pos, vel, m[:, None]], axis=1).astype(np.float64).tofile(f)
    return pos, vel, m


def args(fin, fout, k, b, periodic, rball):
    return [str(fin), str(fout), str(k), repr(b), str(periodic), "1.0", repr(rball)]


def read_output(path):
    with open(path, "rb") as f:
        n, k = np.fromfile(f, np.int32, 2)
        out = {"ids": np.fromfile(f, np.int32, n * k).reshape(n, k), "rho": np.fromfile(f, np.float64, n),
               "lab": np.fromfile(f, np.int32, n), "lab_check": np.fromfile(f, np.int32, n)}
        cnt = np.fromfile(f, np.int32, n)
        flat = np.fromfile(f, np.int32, int(cnt.sum()))
        off = np.concatenate([[0], np.cumsum(cnt)])
        out["ball"] = [np.sort(flat[off[i]:off[i + 1]]) for i in range(n)]
        out["ptnn"] = np.fromfile(f, np.int32, 8 * k).reshape(8, k)
        out["flags"] = np.fromfile(f, np.int32, 8)
        out["lhead"] = np.fromfile(f, np.int32, n)
    return out


def check(out, pos, m, k, b, periodic, rball):
    n = pos.shape[0]
    fl = out["flags"]
    assert fl[0] == 0, "per-particle FindNearest differs from the batch"
    assert fl[1] == 0, "particle fields lost in the reorder"
    assert fl[2] == 0, "input order not restored by ~KDTree"
    assert fl[4] == 0, "SearchBallPosTagged array and vector forms differ"
    assert fl[5] == 0, "FOF linked lists malformed"
    per = 1.0 if periodic else None
    ot = oracle.Tree(pos, mass=m, b=16, kerntype=oracle.KSPH, period=per)
    oi, _ = ot.knn(k, 0)
    assert np.array_equal(np.sort(out["ids"], 1), np.sort(oi, 1)), "FindNearest neighbour sets"
    ro = ot.calc_density(k)
    assert np.max(np.abs(out["rho"] / ro - 1)) < 1e-12, "CalcDensity"
    ol, og = oracle.fof_brute(pos, b, 8, period=per)
    assert fl[3] == og and np.array_equal(out["lab"], ol), "FOF labels"
    # the linked lists partition the particles into every linked component,
    # the groups pruned below minnum included (KDFOF.cxx:98-104)
    al, ag = oracle.fof_brute(pos, b, 1, period=per)
    hl = out["lhead"]
    assert np.array_equal(al[hl], al), "FOF list heads outside their component"
    assert len(np.unique(hl)) == ag, "FOF lists: one per linked component"
    sz = np.bincount(al)[1:]
    assert ((sz > 1) & (sz < 8)).any(), "the case has pruned components of several members"
    # FOF with every fifth particle left out (check = -1 -> label 0)
    keep = np.arange(n) % 5 != 0
    sl, sg = oracle.fof_brute(pos[keep], b, 8, period=per)
    want = np.zeros(n, dtype=np.int32)
    want[keep] = sl
    assert fl[6] == sg and np.array_equal(out["lab_check"], want), "FOF with a check function"
    # SearchBallPosTagged(i, r2): every particle with d2 < r2; the target
    # itself left out on a non-periodic tree, kept on a periodic one
    r2 = rball * rball
    for i in range(n):
        g = np.full(n, -1, dtype=np.int32)
        d2 = np.zeros(n)
        oracle.ball_mark(pos, pos[i], r2, 1, g, d2, exclude=-1 if periodic else i, period=per)
        assert np.array_equal(out["ball"][i], np.flatnonzero(g == 1)), f"SearchBallPosTagged row {i}"
    # FindNearest(x) at 8 points near particles (period-aware)
    q = np.array([pos[p * 97 % n] + [1e-3, 0, 0] for p in range(8)])
    pi, _ = ot.knn_points(q, k)
    assert np.array_equal(np.sort(out["ptnn"], 1), np.sort(pi, 1)), "FindNearest(x)"

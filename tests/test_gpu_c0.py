"""GPU parity at the small CPU configuration's full size and for the periodic
phase-space searches.

C0: N = 32^3 = 32768 uniform particles in the periodic unit box, b = 16, k = 32,
FOF at 0.2 x the mean spacing (0.00625) with minnum 8 (and minnum 2, where
groups are many).  Periodic phase space: FindNearestPhase(tt) and
FindNearestPhase(x, v) search the primary point and all 7 spatial images
(KDSplitNode.cxx:1153-1179).  Bar: neighbour sets, d2 and canonical FOF labels
bit-identical to the oracle.
"""
import numpy as np
import pytest

import oracle
from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu
C0_N, C0_SEED = 32768, 20261202


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c0():
    pos = datagen.uniform(C0_N, 1.0, C0_SEED)
    return pos, oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=1.0)


@pytest.mark.parametrize("variant", [0, 1])
def test_c0_periodic_findnearest_full(gpu, c0, variant):
    """Periodic FindNearest(tt) (k+1 search, smallest dropped) and
    FindNearestPos(tt) (largest dropped), with the reference's image rule
    including its squared-distance-against-distance test (KDSplitNode.cxx:1119-1148)."""
    import nbodylib_amd as nb

    pos, ot = c0
    say(f"C0 periodic FindNearest variant={variant} n={C0_N} k=32")
    t = nb.KDTree(pos, bucket_size=16, kerntype=nb.KSPH, period=1.0)
    nn, d2 = t._knn(32, variant)
    oi, od = ot.knn(32, variant)
    assert np.array_equal(np.sort(nn, 1), np.sort(oi, 1))
    assert np.array_equal(d2, od)


@pytest.mark.parametrize("minnum", [8, 2])
def test_c0_fof_full(gpu, c0, minnum):
    import nbodylib_amd as nb

    pos, ot = c0
    b = 0.2 * datagen.mean_spacing(C0_N, 1.0)
    assert abs(b - 0.00625) < 1e-15
    say(f"C0 FOF n={C0_N} b={b} minnum={minnum}")
    t = nb.KDTree(pos, bucket_size=16, kerntype=nb.KSPH, period=1.0)
    lab, ng = t.FOF(b, minnum)
    raw, _ = ot.fof(b, minnum)
    ol, og = oracle.fof_canonical(raw, minnum)
    assert ng == og and np.array_equal(lab, ol)
    if minnum == 2:
        assert ng > 100


def _periodic_phase_set(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)), rng.normal(0.0, 0.04, (n, 3))


@pytest.mark.parametrize("k", [32, 64])
def test_phase_periodic_knn_vs_brute(gpu, k):
    import nbodylib_amd as nb

    n = 6000
    say(f"periodic phase FindNearestPhase(tt) n={n} k={k}")
    x, v = _periodic_phase_set(n, 22)
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, period=1.0)
    nn, d2 = t.FindNearestPhase(k)
    bi, bd = oracle.knn_brute(x, k, mode=1, period=1.0, vel=v, nd=6)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


def _phase_points_brute(x, v, qx, qv, k, period):
    """k nearest in 6-D from each phase point, over the primary point and the
    images of its position shifted toward the far wall on every subset of the
    axes (all taken), nothing excluded; ascending (d2, ID)."""
    shifts = [()]
    if period is not None:
        shifts = [s for r in range(4) for s in __import__("itertools").combinations(range(3), r)]
    ph = np.concatenate([x, v], axis=1)
    out_i, out_d = [], []
    for q in range(qx.shape[0]):
        best = np.full(ph.shape[0], np.inf)
        for s in shifts:
            p = qx[q].copy()
            for a in s:
                p[a] = p[a] + period if qx[q, a] < period / 2.0 else p[a] - period
            pt = np.concatenate([p, qv[q]])
            d = ph - pt
            d2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) + d[:, 3] * d[:, 3]
            d2 = (d2 + d[:, 4] * d[:, 4]) + d[:, 5] * d[:, 5]
            best = np.minimum(best, d2)
        o = np.lexsort((np.arange(ph.shape[0]), best))[:k]
        out_i.append(o)
        out_d.append(best[o])
    return np.array(out_i, dtype=np.int32), np.array(out_d)


@pytest.mark.parametrize("periodic", [False, True])
def test_phase_points_vs_brute(gpu, periodic):
    """FindNearestPhase(x, v) at arbitrary phase points (KDFindNearest.cxx:528-538)."""
    import nbodylib_amd as nb

    n, m, k = 5000, 200, 32
    say(f"phase points periodic={periodic} n={n} m={m} k={k}")
    x, v = _periodic_phase_set(n, 23)
    rng = np.random.default_rng(24)
    qx, qv = rng.random((m, 3)), rng.normal(0.0, 0.04, (m, 3))
    qx[:20] = x[:20]  # some points exactly at particles (d2 = 0 kept)
    qv[:20] = v[:20]
    per = 1.0 if periodic else None
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, period=per)
    nn, d2 = t.FindNearestPhasePoints(qx, qv, k)
    bi, bd = _phase_points_brute(x, v, qx, qv, k, per)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)
    assert np.all(d2[:20, 0] == 0.0)

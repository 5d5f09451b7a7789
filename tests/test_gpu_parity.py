"""GPU parity of the plain (FP64) form against the CPU oracle.

Each case prints its name, sizes and seed before touching the GPU.
Bar: neighbour ID sets and FOF labels bit-identical, kNN squared distances
bit-identical, densities within 1e-12 relative.
"""
import numpy as np
import pytest

import oracle
from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu
DENS_TOL = 1e-12


def _nb():
    import nbodylib_amd as nb

    return nb


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n,k,seed", [(4096, 32, 1), (5000, 16, 2), (777, 32, 3)])
def test_knn_nonperiodic_vs_brute(gpu, n, k, seed):
    nb = _nb()
    say(f"knn nonperiodic n={n} k={k} seed={seed}")
    pos = datagen.uniform(n, 1.0, seed)
    t = nb.KDTree(pos, kerntype=nb.KSPH)
    nn, d2 = t.FindNearest(k)
    bi, bd = oracle.knn_brute(pos, k, mode=0)
    assert np.array_equal(nn, bi)
    assert np.array_equal(d2, bd)


@pytest.mark.parametrize("variant,mode", [(0, 1), (1, 2)])
def test_knn_periodic_vs_brute(gpu, variant, mode):
    nb = _nb()
    n, k = 4096, 32
    say(f"knn periodic variant={variant} n={n} k={k}")
    pos = datagen.uniform(n, 1.0, 11)
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=1.0)
    nn, d2 = t._knn(k, variant)
    bi, bd = oracle.knn_brute(pos, k, mode=mode, period=1.0)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


def test_density_c0(gpu):
    """C0: N=32^3 uniform periodic unit box, b=16, k=32, KSPH."""
    nb = _nb()
    n, k = 32768, 32
    say(f"density C0 n={n} k={k}")
    pos = datagen.uniform(n, 1.0, 20261202)
    t = nb.KDTree(pos, bucket_size=16, kerntype=nb.KSPH, period=1.0)
    rho, h = t.CalcDensity(k, return_h=True)
    ot = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=1.0)
    ro = ot.calc_density(k)
    assert np.max(np.abs(rho / ro - 1.0)) < DENS_TOL
    ids, d2 = ot.knn(k, 3)
    assert np.allclose(h, 0.5 * np.sqrt(d2[:, -1]), rtol=0, atol=0)
    nn, _ = t.FindNearestNonPeriodic(k)
    assert np.array_equal(np.sort(nn, 1), np.sort(ids, 1))


def test_density_gather(gpu):
    nb = _nb()
    n, k = 6000, 32
    say(f"density gather n={n} k={k}")
    pos = datagen.uniform(n, 1.0, 5)
    m = datagen.lognormal_masses(n, 5)
    t = nb.KDTree(pos, mass=m, kerntype=nb.KSPH)
    rho, h = t.CalcDensityParticleAll(k)
    ot = oracle.Tree(pos, mass=m, b=16, kerntype=oracle.KSPH)
    ro, ho = ot.calc_density_gather(k)
    assert np.max(np.abs(rho / ro - 1.0)) < DENS_TOL
    assert np.array_equal(h, ho)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_density_position(gpu, form):
    """CalcDensityPosition at positions inside and outside the box of a
    periodic clustered tree (the search ignores the period), at particle
    positions (the particle is its own neighbour) and for k = 1, 2, 64."""
    nb = _nb()
    n = 6000
    say(f"density position form={form} n={n}")
    pos, _ = datagen.clustered(n, 1.0, nhalo=6, seed=33)
    m = datagen.lognormal_masses(n, 33)
    rng = np.random.default_rng(33)
    x = np.concatenate([rng.uniform(-0.2, 1.2, (500, 3)), pos[:300], pos[300:400] + 1e-4])
    fm = nb.FORM_PLAIN if form == "plain" else nb.FORM_MIXED
    t = nb.KDTree(pos, mass=m, kerntype=nb.KSPH, period=1.0, form=fm)
    ot = oracle.Tree(pos, mass=m, b=16, kerntype=oracle.KSPH, period=1.0)
    for k in (32, 1, 2, 64):
        # k = 1 at a particle's own position gives h = 0 (0/0 in the reference)
        xs = x if k > 1 else np.concatenate([x[:500], x[800:]])
        say(f"  k={k} m={len(xs)}")
        rho, h = t.CalcDensityPosition(xs, k, return_h=True)
        ro, ho = ot.calc_density_position(xs, k)
        assert np.array_equal(h, ho)
        ok = ro != 0  # k = 1: the one neighbour sits at q = 2, where W = 0
        assert np.all(rho[~ok] == 0)
        assert not ok.any() or np.max(np.abs(rho[ok] / ro[ok] - 1.0)) < DENS_TOL
    one = t.CalcDensityPosition(x[7], 32)
    assert abs(one / ot.calc_density_position(x[7:8], 32)[0][0] - 1.0) < DENS_TOL
    small = nb.KDTree(pos[:20], mass=m[:20], kerntype=nb.KSPH)
    with pytest.raises(Exception):
        small.CalcDensityPosition(x[:2], 32)


@pytest.mark.parametrize("periodic", [False, True])
def test_fof_vs_brute(gpu, periodic):
    nb = _nb()
    n = 8000
    pos = datagen.uniform(n, 1.0, 7)
    b = 0.5 * datagen.mean_spacing(n, 1.0)
    say(f"fof n={n} b={b} periodic={periodic}")
    per = 1.0 if periodic else None
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per)
    lab, ng = t.FOF(b, 8)
    ol, og = oracle.fof_brute(pos, b, 8, period=per)
    assert ng == og
    assert np.array_equal(lab, ol)


@pytest.mark.parametrize("periodic", [False, True])
def test_fof6d_vs_brute(gpu, periodic):
    # FOF on a phase-space (TPHS) tree: 6-D links, d2 = |dx|^2 + |dv|^2
    nb = _nb()
    n = 6000
    pos = datagen.uniform(n, 1.0, 31)
    vel = np.random.default_rng(32).random((n, 3)) * 0.3
    b = 0.85 * (0.3 ** 3 / n) ** (1 / 6)
    say(f"fof6d n={n} b={b} periodic={periodic}")
    per = 1.0 if periodic else None
    t = nb.KDTree(pos, vel=vel, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, period=per)
    lab, ng = t.FOF(b, 8)
    ol, og = oracle.fof_brute(pos, b, 8, period=per, vel=vel)
    assert og > 5
    assert ng == og
    assert np.array_equal(lab, ol)


@pytest.mark.parametrize("periodic", [False, True])
def test_fof_criterion_3d(gpu, periodic):
    """FOFCriterion with FOF3d on a TPHYS tree: the comparator's closure,
    labels bit-identical; a wider search scale changes nothing."""
    nb = _nb()
    n = 6000
    pos = datagen.uniform(n, 1.0, 35)
    per = 1.0 if periodic else None
    b2 = (0.6 * datagen.mean_spacing(n, 1.0)) ** 2
    say(f"fof criterion 3d n={n} periodic={periodic}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per)
    prm = [nb.TPHYS, b2, 0, 0, 0, 0, b2]
    lab, ng = t.FOFCriterion(nb.FOF_CMP_3D, prm, 8)
    ol, og = oracle.fof_criterion_brute(pos, 1, prm, 8, period=per)
    assert og > 5 and ng == og and np.array_equal(lab, ol)
    lab2, _ = t.FOFCriterion(nb.FOF_CMP_3D, [nb.TPHYS, 4 * b2, 0, 0, 0, 0, b2], 8)
    assert np.array_equal(lab2, lab)


@pytest.mark.parametrize("periodic", [False, True])
def test_fof_criterion_6d(gpu, periodic):
    """FOFCriterion with FOF6d on a TPHS tree, unequal position and velocity
    scales: labels bit-identical to the comparator's closure."""
    nb = _nb()
    n = 6000
    pos = datagen.uniform(n, 1.0, 36)
    vel = np.random.default_rng(37).random((n, 3)) * 0.3
    per = 1.0 if periodic else None
    b6 = 0.85 * (0.3 ** 3 / n) ** (1 / 6)
    sx, sv = 0.7 * b6 * b6, 1.5 * b6 * b6
    say(f"fof criterion 6d n={n} periodic={periodic}")
    t = nb.KDTree(pos, vel=vel, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, period=per)
    prm = [nb.TPHS, sx, sv, 0, 0, 0, sx, sv]
    lab, ng = t.FOFCriterion(nb.FOF_CMP_6D, prm, 8)
    ol, og = oracle.fof_criterion_brute(pos, 2, prm, 8, period=per, vel=vel)
    assert og > 5 and ng == og and np.array_equal(lab, ol)


def test_fof_criterion_clustered_and_rejected(gpu):
    """FOF6d on clustered phase-space data (halo cores link through both
    terms), then the cases the call reports as unsupported: a comparator on
    the wrong tree type, a search scale narrower than the comparator's, and
    an unknown comparator."""
    nb = _nb()
    n = 12000
    pos, vel, _ = datagen.clustered(n, 1.0, nhalo=12, seed=38, with_vel=True)
    vel = vel / vel.std()
    sx, sv = (0.3 * datagen.mean_spacing(n, 1.0)) ** 2, 0.05
    say(f"fof criterion 6d clustered n={n}")
    t6 = nb.KDTree(pos, vel=vel, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    prm = [nb.TPHS, sx, sv, 0, 0, 0, sx, sv]
    lab, ng = t6.FOFCriterion(nb.FOF_CMP_6D, prm, 8)
    ol, og = oracle.fof_criterion_brute(pos, 2, prm, 8, vel=vel)
    assert og > 5 and ng == og and np.array_equal(lab, ol)
    t3 = nb.KDTree(pos, kerntype=nb.KSPH)
    for tree, cmp, q in ((t3, nb.FOF_CMP_6D, prm), (t6, nb.FOF_CMP_3D, [nb.TPHYS, sx, 0, 0, 0, 0, sx]),
                         (t6, nb.FOF_CMP_6D, [nb.TPHS, 0.5 * sx, sv, 0, 0, 0, sx, sv]),
                         (t3, 7, [nb.TPHYS, sx, 0, 0, 0, 0, sx])):
        with pytest.raises(nb.NbxError):
            tree.FOFCriterion(cmp, q, 8)


def test_fof6d_clustered_vs_tree_oracle(gpu):
    nb = _nb()
    n = 32768
    pos, _ = datagen.clustered(n, 1.0, nhalo=40, seed=33)
    vel = np.random.default_rng(34).normal(size=(n, 3)) * 0.02
    b = 0.25 * datagen.mean_spacing(n, 1.0)
    say(f"fof6d clustered n={n} b={b}")
    t = nb.KDTree(pos, vel=vel, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    lab, ng = t.FOF(b, 8)
    ot = oracle.Tree(pos, vel=vel, b=16, treetype=oracle.TPHS, kerntype=oracle.KSPH)
    raw, og = ot.fof(b, 8)
    can, _ = oracle.fof_canonical(raw, 8)
    assert og > 0
    assert ng == og
    assert np.array_equal(lab, can)


def test_fof_clustered_vs_tree_oracle(gpu):
    nb = _nb()
    n = 65536
    pos, _ = datagen.clustered(n, 100.0, nhalo=50, seed=3)
    b = 0.2 * datagen.mean_spacing(n, 100.0)
    say(f"fof clustered n={n} b={b}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=100.0)
    lab, ng = t.FOF(b, 8)
    ot = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=100.0)
    raw, og = ot.fof(b, 8)
    can, _ = oracle.fof_canonical(raw, 8)
    assert ng == og
    assert np.array_equal(lab, can)


def test_clustered_knn_density(gpu):
    nb = _nb()
    n, k = 65536, 32
    pos, _ = datagen.clustered(n, 100.0, nhalo=50, seed=4)
    say(f"clustered knn+density n={n} k={k}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=100.0)
    rho = t.CalcDensity(k)
    ot = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=100.0)
    ro = ot.calc_density(k)
    assert np.max(np.abs(rho / ro - 1.0)) < DENS_TOL
    nn, d2 = t.FindNearest(k)
    oi, od = ot.knn(k, 0)
    assert np.array_equal(np.sort(nn, 1), np.sort(oi, 1))
    assert np.array_equal(d2, od)


@pytest.mark.parametrize("periodic", [False, True])
def test_ball_mark(gpu, periodic):
    # SearchBallPos(x | tt, fdist2, imark, nn, dist2): marks and d2 by ID
    nb = _nb()
    n = 20000
    pos = datagen.uniform(n, 1.0, 43)
    per = 1.0 if periodic else None
    say(f"ball mark n={n} periodic={periodic}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per)
    for x, excl in (([0.5, 0.4, 0.6], -1), ([0.02, 0.98, 0.03], -1), (pos[123], 123)):
        g = np.zeros(n, dtype=np.int32)
        d2 = np.full(n, -1.0)
        t.SearchBallPos(np.asarray(x), 0.004, 5, g, d2, exclude_id=excl)
        og = np.zeros(n, dtype=np.int32)
        od2 = np.full(n, -1.0)
        oracle.ball_mark(pos, x, 0.004, 5, og, od2, exclude=excl, period=per)
        assert (og == 5).sum() > 2
        assert np.array_equal(g, og)
        assert np.array_equal(d2, od2)


def test_ball_search(gpu):
    nb = _nb()
    n = 5000
    pos = datagen.uniform(n, 1.0, 9)
    say(f"ball n={n}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=1.0)
    for x in ([0.5, 0.5, 0.5], [0.01, 0.99, 0.02]):
        got = np.sort(t.SearchBallPosTagged(np.array(x), 0.01))
        # brute force with the reference's image rule (images when r2 > s^2)
        exp = []
        xs = [np.array(x)]
        x0 = np.array(x)
        for ax in ([0], [1], [2], [0, 1], [0, 2], [1, 2], [0, 1, 2]):
            xp = x0.copy()
            s = 0.0
            for a in ax:
                if x0[a] < 0.5:
                    xp[a] = x0[a] + 1.0
                    s += x0[a] * x0[a]
                else:
                    xp[a] = x0[a] - 1.0
                    s += xp[a] * xp[a]
            s = np.sqrt(s) if len(ax) > 1 else (x0[ax[0]] if x0[ax[0]] < 0.5 else -xp[ax[0]])
            if 0.01 > s * s:
                xs.append(xp)
        for q in xs:
            d = q[None, :] - pos
            d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            exp.extend(np.nonzero(d2 < 0.01)[0].tolist())
        assert np.array_equal(got, np.sort(np.array(exp, dtype=np.int32)))


def _ball_brute(pos, x, r2, period):
    """IDs with d2 < r2 around x and its images (image when r2 > s^2, the
    position shifted first: SearchBallPosPeriodicTagged), with their d2."""
    pts = [np.array(x, dtype=np.float64)]
    if period is not None:
        x0 = pts[0]
        for ax in ([0], [1], [2], [0, 1], [0, 2], [1, 2], [0, 1, 2]):
            xp = x0.copy()
            s = 0.0
            for a in ax:
                if x0[a] < period / 2.0:
                    xp[a] = x0[a] + period
                    s += x0[a] * x0[a]
                else:
                    xp[a] = x0[a] - period
                    s += xp[a] * xp[a]
            s = np.sqrt(s) if len(ax) > 1 else (x0[ax[0]] if x0[ax[0]] < period / 2.0 else -xp[ax[0]])
            if r2 > s * s:
                pts.append(xp)
    ids, d2s = [], []
    for q in pts:
        d = q[None, :] - pos
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        hit = np.nonzero(d2 < r2)[0]
        ids.extend(hit.tolist())
        d2s.extend(d2[hit].tolist())
    return np.array(ids, dtype=np.int32), np.array(d2s)


@pytest.mark.parametrize("periodic", [False, True])
def test_ball_search_many(gpu, periodic):
    """SearchBallPosTagged for 700 positions in one call (inside, outside and
    on the box edges, and on particles): every row's IDs and d2 equal a
    brute force with the reference's image rule; an empty radius gives empty
    rows; too small a cap writes only the offsets."""
    nb = _nb()
    n, m = 5000, 700
    pos = datagen.uniform(n, 1.0, 46)
    rng = np.random.default_rng(46)
    x = np.concatenate([rng.uniform(-0.1, 1.1, (400, 3)), pos[:200], rng.uniform(0, 1, (100, 3)) * [1, 0, 1]])
    per = 1.0 if periodic else None
    say(f"ball many n={n} m={m} periodic={periodic}")
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per)
    r2 = 0.006
    off, ids, d2 = t.SearchBallPosTaggedMany(x, r2)
    assert off[0] == 0 and np.all(np.diff(off) >= 0)
    for i in range(m):
        bi, bd = _ball_brute(pos, x[i], r2, per)
        row, rd = ids[off[i]:off[i + 1]], d2[off[i]:off[i + 1]]
        o1, o2 = np.argsort(row, kind="stable"), np.argsort(bi, kind="stable")
        assert np.array_equal(row[o1], bi[o2]), i
        assert np.array_equal(rd[o1], bd[o2]), i
    off0, ids0, _ = t.SearchBallPosTaggedMany(x, 0.0)
    assert off0[-1] == 0 and ids0.size == 0
    o = np.zeros(m + 1, dtype=np.int64)
    got = np.full(4, -7, dtype=np.int32)
    qq = np.ascontiguousarray(x.T)
    nb._check(nb.lib().nbx_ball_points_many(t._h, nb._p(qq), m, r2, nb._p(o), nb._p(got), None, 3))
    assert np.array_equal(o, off) and np.all(got == -7)

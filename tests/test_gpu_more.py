"""GPU parity: 6-D phase space, velocity density, velocity (TVEL) trees,
position queries and the differentiation form (adjoint)."""
import numpy as np
import pytest

import oracle
from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _phase_set(n, seed):
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=max(4, n // 2000), seed=seed, with_vel=True)
    return pos / pos.std(0), vel / vel.std(0)


@pytest.mark.parametrize("k", [32, 64])
def test_phase_knn_vs_brute(gpu, k):
    import nbodylib_amd as nb

    n = 4000
    say(f"phase knn n={n} k={k}")
    x, v = _phase_set(n, 31)
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    nn, d2 = t.FindNearestPhase(k)
    bi, bd = oracle.knn_brute(x, k, mode=0, vel=v, nd=6)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


def test_phase_density_vs_oracle(gpu):
    """TPHS CalcDensity with the metric factor taken as 1 (6-D norm 3024/(762 pi^3))."""
    import nbodylib_amd as nb

    n, k = 20000, 64
    say(f"phase density n={n} k={k}")
    x, v = _phase_set(n, 32)
    m = datagen.lognormal_masses(n, 32)
    t = nb.KDTree(x, mass=m, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    rho = t.CalcDensity(k)
    ro = oracle.Tree(x, mass=m, vel=v, b=16, treetype=oracle.TPHS, kerntype=oracle.KSPH).calc_density(k)
    assert np.max(np.abs(rho / ro - 1.0)) < 1e-12


def test_vel_density_vs_oracle(gpu):
    import nbodylib_amd as nb

    n = 20000
    say(f"vel density n={n}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=33, with_vel=True)
    t = nb.KDTree(pos, vel=vel, kerntype=nb.KSPH)
    rv = t.CalcVelDensity(32, 64)
    ro = oracle.Tree(pos, vel=vel, b=16, kerntype=oracle.KSPH).calc_veldensity(32, 64)
    assert np.max(np.abs(rv / ro - 1.0)) < 1e-12


def test_veldensity_position_vs_oracle(gpu):
    """CalcVelDensityPosition at random phase points and at the particles' own
    (x, v), against the oracle; Nsmooth > Nsearch clamps; a tree without
    velocities and too few particles are errors."""
    import nbodylib_amd as nb

    n, m = 20000, 3000
    say(f"vel density position n={n} m={m}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=38, with_vel=True)
    rng = np.random.default_rng(38)
    x = np.concatenate([rng.uniform(0.0, 100.0, (m, 3)), pos[:500]])
    v = np.concatenate([rng.normal(0.0, 300.0, (m, 3)), vel[:500]])
    t = nb.KDTree(pos, vel=vel, kerntype=nb.KSPH, period=100.0)
    ot = oracle.Tree(pos, vel=vel, b=16, kerntype=oracle.KSPH, period=100.0)
    ref = {}
    for ns, nS in ((32, 64), (16, 16), (64, 32)):
        rho, hv = t.CalcVelDensityPosition(x, v, ns, nS, return_h=True)
        ro, ho = ot.calc_veldensity_position(x, v, ns, nS)
        assert np.max(np.abs(hv / ho - 1.0)) < 1e-14
        assert np.max(np.abs(rho / ro - 1.0)) < 1e-12
        ref[(ns, nS)] = rho
    one = t.CalcVelDensityPosition(x[7], v[7], 32, 64)
    assert one == ref[(32, 64)][7]
    with pytest.raises(nb.NbxError):
        nb.KDTree(pos, kerntype=nb.KSPH).CalcVelDensityPosition(x[:2], v[:2], 8, 16)
    with pytest.raises(nb.NbxError):
        nb.KDTree(pos[:10], vel=vel[:10], kerntype=nb.KSPH).CalcVelDensityPosition(x[:2], v[:2], 8, 16)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_veldensity_phys_vs_oracle(gpu, form):
    """CalcVelDensityWithPhysDensity over the tree's own CalcDensity and over
    given densities, against the oracle (zeros, where every neighbour lies
    beyond r = 2h, compared as zeros)."""
    import nbodylib_amd as nb

    n = 20000
    say(f"vel density phys n={n} form={form}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=40, with_vel=True)
    mass = datagen.lognormal_masses(n, 40)
    f = nb.FORM_PLAIN if form == "plain" else nb.FORM_MIXED
    t = nb.KDTree(pos, mass=mass, vel=vel, kerntype=nb.KSPH, form=f)
    ot = oracle.Tree(pos, mass=mass, vel=vel, b=16, kerntype=oracle.KSPH)
    given = datagen.lognormal_masses(n, 41) * 50.0
    for ns, nS, rho in ((32, 64, None), (16, 48, given)):
        got = t.CalcVelDensityWithPhysDensity(ns, nS, rho)
        ro = ot.calc_density(ns) if rho is None else rho
        want = ot.calc_veldensity_phys(ns, nS, ro)
        z = want == 0
        assert (~z).any()
        assert np.array_equal(got == 0, z)
        assert np.max(np.abs(got[~z] / want[~z] - 1.0)) < 1e-12
    # positions of order 1e-9: r/h far past the table (index beyond 2^31), every term W(2) = 0
    tiny = nb.KDTree(pos * 1e-10, mass=mass, vel=vel, kerntype=nb.KSPH, form=f)
    assert np.array_equal(tiny.CalcVelDensityWithPhysDensity(16, 32, np.ones(n)), np.zeros(n))


@pytest.mark.parametrize("periodic", [False, True])
def test_knn_points_vs_oracle(gpu, periodic):
    import nbodylib_amd as nb

    n, m, k = 5000, 777, 16
    say(f"knn points periodic={periodic}")
    pos = datagen.uniform(n, 1.0, 34)
    q = datagen.uniform(m, 1.0, 35)
    per = 1.0 if periodic else None
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per)
    nn, d2 = t.FindNearestPosPoints(q, k)
    oi, od = oracle.Tree(pos, b=16, period=per).knn_points(q, k)
    assert np.array_equal(np.sort(nn, 1), np.sort(oi, 1))
    assert np.array_equal(d2, od)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_velocity_tree_vs_oracle(gpu, form):
    """TVEL: the tree is built on the velocities and FindNearestVel, the
    velocity-point search and both densities run in velocity space
    (KDFindNearest.cxx:258, :335-346, :530-541; KDCalcSmoothQuantities.cxx:228,
    :783).  Positions are passed too and must be ignored: the oracle here is
    brute force and the oracle's tree, both over the velocities."""
    import nbodylib_amd as nb

    n, m, k = 20000, 500, 32
    say(f"velocity tree n={n} k={k} form={form}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=36, with_vel=True)
    mass = datagen.lognormal_masses(n, 36)
    f = nb.FORM_PLAIN if form == "plain" else nb.FORM_MIXED
    t = nb.KDTree(pos, mass=mass, vel=vel, treetype=nb.TVEL, kerntype=nb.KSPH, form=f)
    nn, d2 = t.FindNearestVel(k)
    bi, bd = oracle.knn_brute(vel, k, mode=0)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)
    qv = vel[:m] + 7.0
    ot = oracle.Tree(vel, mass=mass, b=16, kerntype=oracle.KSPH)
    qi, qd = t.FindNearestVelPoints(qv, k)
    oi, od = ot.knn_points(qv, k)
    assert np.array_equal(np.sort(qi, 1), np.sort(oi, 1))
    assert np.array_equal(qd, od)
    rho = t.CalcDensity(k)
    assert np.max(np.abs(rho / ot.calc_density(k) - 1.0)) < 1e-12
    rg, hg = t.CalcDensityParticleAll(k)
    og, ohg = ot.calc_density_gather(k)
    assert np.array_equal(hg, ohg)
    assert np.max(np.abs(rg / og - 1.0)) < 1e-12
    # calls that need positions are refused on a velocity tree
    for call in (lambda: t.FOF(1.0, 8), lambda: t.FindNearestPos(k)):
        with pytest.raises(nb.NbxError):
            call()
    tp = nb.KDTree(pos, vel=vel, kerntype=nb.KSPH)
    with pytest.raises(nb.NbxError):
        tp.FindNearestVel(k)


def test_adjoint_gradcheck_and_reference(gpu):
    """Differentiation form at N=4096 in FP64: torch.autograd.gradcheck with
    central differences (eps 1e-6) on a subset of coordinates, and the full
    gradient against the oracle's hand-derived FP64 adjoint."""
    import torch

    from nbodylib_amd.device import SPHDensity, sph_density

    n, k = 4096, 32
    say(f"adjoint n={n} k={k}")
    pos = datagen.uniform(n, 1.0, 36)
    mass = datagen.lognormal_masses(n, 36)
    nbr, _ = oracle.knn_brute(pos, k, mode=0)
    dev = torch.device("cuda:0")
    P = torch.tensor(np.ascontiguousarray(pos.T), device=dev, requires_grad=True)
    M = torch.tensor(mass, device=dev, requires_grad=True)
    N = torch.tensor(nbr, device=dev)
    rho = SPHDensity.apply(P, M, N)
    ro, _ = oracle.density_analytic(pos, mass, nbr)
    assert np.max(np.abs(rho.detach().cpu().numpy() / ro - 1)) < 1e-12
    loss = torch.log(rho).sum()
    loss.backward()
    dpos_ref, dm_ref = oracle.density_adjoint(pos, mass, nbr, 1.0 / ro)
    gp = P.grad.cpu().numpy().T
    gm = M.grad.cpu().numpy()
    scale = np.max(np.abs(dpos_ref))
    assert np.max(np.abs(gp - dpos_ref)) < 1e-9 * scale
    assert np.max(np.abs(gm - dm_ref)) < 1e-9 * np.max(np.abs(dm_ref))

    # gradcheck on a small slice: perturb 24 particles' coordinates and masses
    idx = torch.tensor(np.random.default_rng(0).choice(n, 24, replace=False), device=dev)
    base_p = P.detach().clone()
    base_m = M.detach().clone()

    def f(sub_p, sub_m):
        p = base_p.clone()
        p[:, idx] = sub_p
        mm = base_m.clone()
        mm[idx] = sub_m
        return torch.log(SPHDensity.apply(p.contiguous(), mm.contiguous(), N)).sum().reshape(1)

    sp = base_p[:, idx].clone().requires_grad_(True)
    sm = base_m[idx].clone().requires_grad_(True)
    # the adjoint accumulates with FP64 atomics, so repeated backward passes
    # differ in the last bits (summation order).  nondet_tol is absolute and
    # these gradients reach 1e2-1e3, where a few ulps of some 64 terms come
    # to ~1e-11: 1e-9 allows that (1e-12 relative), not more
    assert torch.autograd.gradcheck(f, (sp, sm), eps=1e-6, atol=1e-5, rtol=1e-4, nondet_tol=1e-9)


@pytest.mark.parametrize("form,kind", [("plain", "uniform"), ("mixed", "uniform"), ("plain", "clustered")])
def test_smooth_vel_vs_oracle(gpu, form, kind):
    """CalcSmoothVel and CalcSmoothVelDisp against the oracle's restatement
    (sums in another order: 1e-12 of the largest mean, 1e-11 of the largest
    dispersion entry); densities given (densityset = 1) scale the means
    exactly; the caller's means (meanvelset = 1) reproduce the dispersion."""
    import nbodylib_amd as nb

    n, k = 6000, 32
    say(f"smooth vel form={form} kind={kind} n={n} k={k}")
    if kind == "uniform":
        pos = datagen.uniform(n, 1.0, 43)
        vel = np.random.default_rng(43).normal(size=(n, 3))
    else:
        pos, vel, _ = datagen.clustered(n, 1.0, nhalo=8, seed=44, with_vel=True)
    m = datagen.lognormal_masses(n, 43)
    fm = nb.FORM_PLAIN if form == "plain" else nb.FORM_MIXED
    t = nb.KDTree(pos, mass=m, vel=vel, kerntype=nb.KSPH, form=fm)
    mean, disp = t.CalcSmoothVelDisp(k)
    ot = oracle.Tree(pos, mass=m, vel=vel, b=16, kerntype=oracle.KSPH)
    om, od = ot.smooth_vel(k, disp=True)
    assert np.max(np.abs(mean - om)) < 1e-12 * np.max(np.abs(om))
    assert np.max(np.abs(disp - od)) < 1e-11 * np.max(np.abs(od))
    assert np.array_equal(disp, np.transpose(disp, (0, 2, 1)))
    rho = t.CalcDensity(k)
    m2 = t.CalcSmoothVel(k, rho=2.0 * rho)
    assert np.max(np.abs(2.0 * m2 - om)) < 1e-12 * np.max(np.abs(om))
    _, d2 = t.CalcSmoothVelDisp(k, mean=om)
    assert np.max(np.abs(d2 - od)) < 1e-11 * np.max(np.abs(od))
    osk, oku = ot.smooth_moments(k, om, od)
    sk, ku = t.CalcSmoothVelMoments(k)
    assert np.max(np.abs(sk - osk)) < 1e-9 * np.max(np.abs(osk))
    assert np.max(np.abs(ku - oku)) < 1e-9 * np.max(np.abs(oku))
    sk2 = t.CalcSmoothVelSkew(k, mean=om, disp=od)
    assert np.max(np.abs(sk2 - osk)) < 1e-10 * np.max(np.abs(osk))


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_smooth_local_vs_oracle(gpu, form):
    """CalcSmoothLocalMean/Disp, Particle forms for every particle and
    Position forms at positions inside and outside the box, against the
    oracle (1e-12 of the largest value); given rho = 1 and w = 1 the means
    are the gather densities."""
    import nbodylib_amd as nb

    n, k = 6000, 32
    say(f"smooth local form={form} n={n} k={k}")
    pos, _ = datagen.clustered(n, 1.0, nhalo=8, seed=64)
    m = datagen.lognormal_masses(n, 64)
    w = np.random.default_rng(64).normal(size=n)
    fm = nb.FORM_PLAIN if form == "plain" else nb.FORM_MIXED
    t = nb.KDTree(pos, mass=m, kerntype=nb.KSPH, form=fm)
    ot = oracle.Tree(pos, mass=m, b=16, kerntype=oracle.KSPH)
    mu = t.CalcSmoothLocalMeanAll(w, k)
    om, _ = ot.smooth_local(k, w)
    assert np.max(np.abs(mu - om)) < 1e-12 * np.max(np.abs(om))
    dd = t.CalcSmoothLocalDispAll(w, om, k)
    _, od = ot.smooth_local(k, w, mean_in=om)
    assert np.max(np.abs(dd - od)) < 1e-12 * np.max(np.abs(od))
    x = np.random.default_rng(65).uniform(-0.1, 1.1, (400, 3))
    mp = t.CalcSmoothLocalMeanPosition(x, w, k)
    omp, _ = ot.smooth_local(k, w, q=x)
    assert np.max(np.abs(mp - omp)) < 1e-12 * np.max(np.abs(omp))
    dp = t.CalcSmoothLocalDispPosition(x, w, omp, k)
    _, odp = ot.smooth_local(k, w, mean_in=omp, q=x)
    assert np.max(np.abs(dp - odp)) < 1e-12 * np.max(np.abs(odp))
    g, _ = t.CalcDensityParticleAll(k)
    g1 = t.CalcSmoothLocalMeanAll(np.ones(n), k, rho=np.ones(n))
    assert np.max(np.abs(g1 / g - 1)) < 1e-12

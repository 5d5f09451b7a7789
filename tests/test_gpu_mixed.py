"""GPU parity of the precision form (FP32 search + FP64 re-check band).

Bar: neighbour ID sets identical to the FP64 oracle; h identical (it comes
from the FP64 k-th distance); densities within 2e-6 relative of the FP64
oracle.
"""
import numpy as np
import pytest

import oracle
from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu
MIXED_TOL = 2e-6  # the precision form's stated density bound; FP64 weights give ~1e-15 in practice


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("kind,n,k", [("uniform", 4096, 32), ("uniform", 5000, 16), ("clustered", 65536, 32)])
def test_mixed_knn_sets(gpu, kind, n, k):
    import nbodylib_amd as nb

    say(f"mixed knn {kind} n={n} k={k}")
    if kind == "uniform":
        pos = datagen.uniform(n, 1.0, 21)
        per = None
        ids, _ = oracle.knn_brute(pos, k, mode=0)
    else:
        pos, _ = datagen.clustered(n, 100.0, nhalo=50, seed=22)
        per = 100.0
        ids, _ = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=per).knn(k, 3)
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=per, form=nb.FORM_MIXED)
    nn = np.zeros((n, k), dtype=np.int32)
    nb._check(nb.lib().nbx_knn(t._h, k, nb.KNN_NONPERIODIC, nb.LAYOUT_BY_ID, nb._p(nn), None))
    assert np.array_equal(np.sort(nn, 1), np.sort(ids, 1))
    say(f"  fallbacks={t.last_fallbacks()}")


@pytest.mark.parametrize("kind,n", [("uniform", 32768), ("clustered", 65536)])
def test_mixed_density(gpu, kind, n):
    import nbodylib_amd as nb

    k = 32
    say(f"mixed density {kind} n={n}")
    if kind == "uniform":
        pos, L = datagen.uniform(n, 1.0, 20261202), 1.0
        mass = None
    else:
        pos, _ = datagen.clustered(n, 100.0, nhalo=50, seed=23)
        L = 100.0
        mass = datagen.lognormal_masses(n, 23)
    t = nb.KDTree(pos, mass=mass, kerntype=nb.KSPH, period=L, form=nb.FORM_MIXED)
    rho, h = t.CalcDensity(k, return_h=True)
    ot = oracle.Tree(pos, mass=mass, b=16, kerntype=oracle.KSPH, period=L)
    ro = ot.calc_density(k)
    _, d2 = ot.knn(k, 3)
    rel = np.max(np.abs(rho / ro - 1.0))
    say(f"  max rel {rel:.3e} fallbacks={t.last_fallbacks()}")
    assert rel < MIXED_TOL
    assert np.array_equal(h, 0.5 * np.sqrt(d2[:, -1]))


def test_mixed_duplicates_fall_back(gpu):
    """Exact duplicates (d2 == 0) cannot be decided in FP32.  A single
    duplicate is resolved by the in-kernel FP64 verification; two duplicates
    of one particle send it through the FP64 re-run.  Both must match the
    oracle."""
    import nbodylib_amd as nb

    n, k = 3000, 16
    say(f"mixed duplicates n={n}")
    pos = datagen.uniform(n, 1.0, 24)
    pos[1::7] = pos[0::7][: len(pos[1::7])]
    pos[2::7] = pos[0::7][: len(pos[2::7])]
    t = nb.KDTree(pos, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    rho = t.CalcDensity(k)
    ro = oracle.Tree(pos, b=16, kerntype=oracle.KSPH).calc_density(k)
    assert np.max(np.abs(rho / ro - 1.0)) < MIXED_TOL
    assert t.last_fallbacks() > 0


def _phase_set(n, seed):
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=max(4, n // 2000), seed=seed, with_vel=True)
    return pos / pos.std(0), vel / vel.std(0)


@pytest.mark.parametrize("k", [16, 64])
def test_mixed_phase_knn_sets(gpu, k):
    """6-D phase-space tree, precision form: FindNearestPhase sets equal the
    FP64 brute force."""
    import nbodylib_amd as nb

    n = 6000
    say(f"mixed phase knn n={n} k={k}")
    x, v = _phase_set(n, 34)
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    nn = np.zeros((n, k), dtype=np.int32)
    nb._check(nb.lib().nbx_knn(t._h, k, nb.KNN_FINDNEARESTPHASE, nb.LAYOUT_BY_ID, nb._p(nn), None))
    bi, _ = oracle.knn_brute(x, k, mode=0, vel=v, nd=6)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    say(f"  fallbacks={t.last_fallbacks()}")


@pytest.mark.parametrize("k", [20, 64])
def test_mixed_phase_density(gpu, k):
    """6-D CalcDensity on a TPHS tree in the precision form against the FP64
    oracle (6-D norm 3024/(762 pi^3), metric factor 1), for both kept-key
    widths (k <= 32 and k <= 64); and the FindNearestPhase rows with their
    FP64 distances against brute force."""
    import nbodylib_amd as nb

    n = 20000
    say(f"mixed phase density n={n} k={k}")
    x, v = _phase_set(n, 35)
    m = datagen.lognormal_masses(n, 35)
    t = nb.KDTree(x, mass=m, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    rho = t.CalcDensity(k)
    ro = oracle.Tree(x, mass=m, vel=v, b=16, treetype=oracle.TPHS, kerntype=oracle.KSPH).calc_density(k)
    rel = np.max(np.abs(rho / ro - 1.0))
    say(f"  max rel {rel:.3e} fallbacks={t.last_fallbacks()}")
    assert rel < MIXED_TOL
    nn, d2 = t.FindNearestPhase(k)
    bi, bd = oracle.knn_brute(x, k, mode=0, vel=v, nd=6)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


def test_mixed_sets_full_size():
    """At the headline size (16.8M clustered particles, periodic L = 100,
    k = 32, CalcDensity's search) the precision form's neighbour rows equal
    the FP64 form's exactly, and so do the smoothing lengths; the densities
    agree to the stated bound."""
    import nbodylib_amd as nb
    import torch
    from nbodylib_amd.device import DeviceTree

    n, k, L = 1 << 24, 32, 100.0
    say(f"mixed vs plain sets n={n} k={k}")
    pos, _ = datagen.clustered(n, L, nhalo=2000, seed=20261203)
    dev = torch.device("cuda:0")
    xs = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
    out = {}
    for form in (nb.FORM_PLAIN, nb.FORM_MIXED):
        t = DeviceTree(xs, bucket_size=16, kerntype=nb.KSPH, period=L, form=form)
        nn = torch.empty((n, k), dtype=torch.int32, device=dev)
        rho = torch.empty(n, dtype=torch.float64, device=dev)
        hh = torch.empty(n, dtype=torch.float64, device=dev)
        t.knn_density(k, nn, rho, hh)
        out[form] = (torch.sort(nn, dim=1).values.cpu(), rho.cpu(), hh.cpu(), t.last_fallbacks())
        say(f"  form {form}: fallbacks {out[form][3]}")
        del t, nn
    (np_, rp, hp, _), (nm, rm, hm, _) = out[nb.FORM_PLAIN], out[nb.FORM_MIXED]
    assert torch.equal(np_, nm)
    assert torch.equal(hp, hm)
    assert float(torch.max(torch.abs(rm / rp - 1.0))) < MIXED_TOL

"""GPU parity for the larger searches and the FP64 answers of the precision
form.

* k up to 128 (the reference sizes its queue at run time for any Nsearch,
  KDFindNearest.cxx:250-253): FindNearest non-periodic (k = 129) and
  periodic (k = 128, a k + 1 = 129 search), 6-D phase space, CalcDensity,
  CalcVelDensity with Nsearch = 128 -- against brute force and the oracle.
  On a precision-form tree these run on the FP64 kernel.
* The precision form asked for distances: rows and d2 bit-identical to the
  FP64 form's (FP64 d2 of the selected entries, ascending by (d2, ID)).
* The re-check band: any band no narrower than the proven FP32 error bound
  gives the FP64 form's sets (it only moves work to the FP64 re-check).
"""
import numpy as np
import pytest

import oracle
from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_knn_k129_nonperiodic_vs_brute(gpu, form):
    import nbodylib_amd as nb

    n, k = 6000, 129
    say(f"knn k={k} nonperiodic n={n} form={form}")
    pos = datagen.uniform(n, 1.0, 81)
    f = nb.FORM_MIXED if form == "mixed" else nb.FORM_PLAIN
    t = nb.KDTree(pos, kerntype=nb.KSPH, form=f)
    nn, d2 = t.FindNearest(k)
    bi, bd = oracle.knn_brute(pos, k, mode=0)
    assert np.array_equal(nn, bi)
    assert np.array_equal(d2, bd)


@pytest.mark.parametrize("variant,mode", [(0, 1), (1, 2)])
def test_knn_k128_periodic_vs_brute(gpu, variant, mode):
    import nbodylib_amd as nb

    n, k = 6000, 128
    say(f"knn k={k} periodic variant={variant} n={n}")
    pos = datagen.uniform(n, 1.0, 82)
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=1.0)
    nn, d2 = t._knn(k, variant)
    bi, bd = oracle.knn_brute(pos, k, mode=mode, period=1.0)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


def test_phase_knn_k100_vs_brute(gpu):
    import nbodylib_amd as nb

    n, k = 5000, 100
    say(f"phase knn k={k} n={n}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=4, seed=83, with_vel=True)
    x, v = pos / pos.std(0), vel / vel.std(0)
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    nn, d2 = t.FindNearestPhase(k)
    bi, bd = oracle.knn_brute(x, k, mode=0, vel=v, nd=6)
    assert np.array_equal(np.sort(nn, 1), np.sort(bi, 1))
    assert np.array_equal(d2, bd)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_density_k128_vs_oracle(gpu, form):
    import nbodylib_amd as nb

    n, k = 20000, 128
    say(f"CalcDensity k={k} n={n} form={form}")
    pos = datagen.clustered(n, 100.0, nhalo=20, seed=84)[0]
    m = datagen.lognormal_masses(n, 84)
    f = nb.FORM_MIXED if form == "mixed" else nb.FORM_PLAIN
    t = nb.KDTree(pos, mass=m, kerntype=nb.KSPH, period=100.0, form=f)
    rho, h = t.CalcDensity(k, return_h=True)
    ot = oracle.Tree(pos, mass=m, b=16, kerntype=oracle.KSPH, period=100.0)
    ro = ot.calc_density(k)
    _, od2 = ot.knn(k, 3)
    assert np.max(np.abs(rho / ro - 1.0)) < 1e-12
    assert np.array_equal(h, 0.5 * np.sqrt(od2[:, -1]))


def test_veldensity_nsearch128_vs_oracle(gpu):
    import nbodylib_amd as nb

    n = 20000
    say(f"CalcVelDensity(64, 128) n={n}")
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=85, with_vel=True)
    t = nb.KDTree(pos, vel=vel, kerntype=nb.KSPH)
    rv = t.CalcVelDensity(64, 128)
    ro = oracle.Tree(pos, vel=vel, b=16, kerntype=oracle.KSPH).calc_veldensity(64, 128)
    assert np.max(np.abs(rv / ro - 1.0)) < 1e-12


@pytest.mark.parametrize("k,layout", [(32, 0), (32, 1), (64, 0)])
def test_mixed_fp64_distances_equal_plain(gpu, k, layout):
    """nbx_knn with d2 on a precision-form tree: the FP32 search, then each
    row's FP64 d2 and (d2, ID) order; identical to the FP64 form's rows."""
    import nbodylib_amd as nb

    n = 65536
    say(f"mixed form with d2 k={k} layout={layout} n={n}")
    pos = datagen.clustered(n, 100.0, nhalo=40, seed=86)[0]
    out = {}
    for form in (nb.FORM_PLAIN, nb.FORM_MIXED):
        t = nb.KDTree(pos, kerntype=nb.KSPH, form=form)
        out[form] = t._knn(k, nb.KNN_FINDNEAREST, layout)
        if layout == 1:
            out[form] = (t.order(), out[form])
    if layout == 1:
        # slot spaces are the same tree order for both forms
        assert np.array_equal(out[nb.FORM_PLAIN][0], out[nb.FORM_MIXED][0])
        out = {f: v[1] for f, v in out.items()}
    assert np.array_equal(out[nb.FORM_MIXED][0], out[nb.FORM_PLAIN][0])
    assert np.array_equal(out[nb.FORM_MIXED][1], out[nb.FORM_PLAIN][1])


@pytest.mark.parametrize("k", [1, 7, 20, 33, 50])
def test_mixed_rows_odd_k(gpu, k):
    """Precision-form rows for k that do not divide the 64 lanes of a wave
    (the output rows of a wave wrap across lanes), against the FP64 form's,
    by ID and by slot, with distances."""
    import nbodylib_amd as nb

    n = 20011
    say(f"mixed rows odd k={k} n={n}")
    pos = datagen.clustered(n, 100.0, nhalo=30, seed=88)[0]
    for layout in (0, 1):
        rows = {}
        for form in (nb.FORM_PLAIN, nb.FORM_MIXED):
            t = nb.KDTree(pos, kerntype=nb.KSPH, form=form)
            rows[form] = t._knn(k, nb.KNN_FINDNEAREST, layout)
        assert np.array_equal(rows[nb.FORM_MIXED][0], rows[nb.FORM_PLAIN][0])
        assert np.array_equal(rows[nb.FORM_MIXED][1], rows[nb.FORM_PLAIN][1])
    tm = nb.KDTree(pos, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    tp = nb.KDTree(pos, kerntype=nb.KSPH, form=nb.FORM_PLAIN)
    rm, rp = tm.CalcDensity(k), tp.CalcDensity(k)
    # (k = 1: the only neighbour sits at r = 2h, where the kernel is 0)
    assert np.all(np.abs(rm - rp) <= 2e-6 * np.abs(rp))


@pytest.mark.parametrize("band", [1e-12, 1e-5, 1e-3])
def test_mixed_recheck_band(gpu, band):
    import nbodylib_amd as nb

    n, k = 65536, 32
    say(f"mixed form recheck_rel={band} n={n} k={k}")
    pos = datagen.clustered(n, 100.0, nhalo=40, seed=87)[0]
    ids, _ = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=100.0).knn(k, 3)
    t = nb.KDTree(pos, kerntype=nb.KSPH, period=100.0, form=nb.FORM_MIXED, recheck_rel=band)
    nn, _ = t.FindNearestNonPeriodic(k)
    assert np.array_equal(np.sort(nn, 1), np.sort(ids, 1))
    say(f"  fallbacks={t.last_fallbacks()}")

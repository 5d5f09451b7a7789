"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle's brute-force restatements) checked against the oracle's tree
path on the CPU and against the HIP path on the GPU.

Bar: neighbour ID sets and squared distances bit-identical, FOF canonical
labels bit-identical, densities within 1e-12 relative (FP64) or 2e-6 (the
precision form).
"""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def _rows_equal(a, b):
    return np.array_equal(np.sort(a, 1), np.sort(b, 1))


def test_golden_uniform_periodic_oracle_tree():
    g = _load("uniform_periodic_n1024.npz")
    t = oracle.Tree(g["pos"], b=16, kerntype=oracle.KSPH, period=float(g["period"]))
    ids, d2 = t.knn(int(g["k"]), 0)
    assert _rows_equal(ids, g["findnearest_ids"]) and np.array_equal(d2, g["findnearest_d2"])
    ids, d2 = t.knn(int(g["k"]), 1)
    assert _rows_equal(ids, g["findnearestpos_ids"]) and np.array_equal(d2, g["findnearestpos_d2"])
    rho = t.calc_density(int(g["density_k"]))
    assert np.max(np.abs(rho / g["rho"] - 1.0)) < 1e-12
    raw, ng = t.fof(float(g["fof_b"]), int(g["fof_minnum"]))
    can, _ = oracle.fof_canonical(raw, int(g["fof_minnum"]))
    assert ng == int(g["fof_ngroups"]) and np.array_equal(can, g["fof_labels"])


def test_golden_clustered_oracle_tree():
    g = _load("clustered_n2048.npz")
    t = oracle.Tree(g["pos"], mass=g["mass"], b=16, kerntype=oracle.KSPH)
    ids, _ = t.knn(int(g["k"]), 0)
    assert _rows_equal(ids, g["ids"])
    rho = t.calc_density(int(g["k"]))
    assert np.max(np.abs(rho / g["rho"] - 1.0)) < 1e-12
    rg, _ = t.calc_density_gather(int(g["k"]))
    assert np.max(np.abs(rg / g["rho_gather"] - 1.0)) < 1e-12
    raw, ng = t.fof(float(g["fof_b"]), int(g["fof_minnum"]))
    can, _ = oracle.fof_canonical(raw, int(g["fof_minnum"]))
    assert ng == int(g["fof_ngroups"]) and np.array_equal(can, g["fof_labels"])


def test_golden_phase6d_oracle_tree():
    g = _load("phase6d_n1024.npz")
    t = oracle.Tree(g["pos"], vel=g["vel"], b=16, treetype=oracle.TPHS, kerntype=oracle.KSPH)
    ids, d2 = t.knn(int(g["k"]), 2)
    assert _rows_equal(ids, g["ids"]) and np.array_equal(d2, g["d2"])


# ---------------------------------------------------------------- GPU
def _nb():
    import nbodylib_amd as nb

    return nb


@pytest.mark.gpu
@pytest.mark.parametrize("form", [0, 1])
def test_golden_uniform_periodic_gpu(gpu, form):
    nb = _nb()
    g = _load("uniform_periodic_n1024.npz")
    print(f"golden uniform periodic n=1024 form={form}", flush=True)
    t = nb.KDTree(g["pos"], kerntype=nb.KSPH, period=float(g["period"]), form=form)
    ids, d2 = t.FindNearest(int(g["k"]))
    assert _rows_equal(ids, g["findnearest_ids"]) and np.array_equal(d2, g["findnearest_d2"])
    ids, d2 = t.FindNearestPos(int(g["k"]))
    assert _rows_equal(ids, g["findnearestpos_ids"]) and np.array_equal(d2, g["findnearestpos_d2"])
    rho, h = t.CalcDensity(int(g["density_k"]), return_h=True)
    tol = 1e-12 if form == 0 else 2e-6
    assert np.max(np.abs(rho / g["rho"] - 1.0)) < tol
    assert np.max(np.abs(h / g["h"] - 1.0)) < 1e-12
    lab, ng = t.FOF(float(g["fof_b"]), int(g["fof_minnum"]))
    assert ng == int(g["fof_ngroups"]) and np.array_equal(lab, g["fof_labels"])


@pytest.mark.gpu
@pytest.mark.parametrize("form", [0, 1])
def test_golden_clustered_gpu(gpu, form):
    nb = _nb()
    g = _load("clustered_n2048.npz")
    print(f"golden clustered n=2048 form={form}", flush=True)
    t = nb.KDTree(g["pos"], mass=g["mass"], kerntype=nb.KSPH, form=form)
    ids, _ = t.FindNearest(int(g["k"]))
    assert _rows_equal(ids, g["ids"])
    tol = 1e-12 if form == 0 else 2e-6
    rho = t.CalcDensity(int(g["k"]))
    assert np.max(np.abs(rho / g["rho"] - 1.0)) < tol
    rg, _ = t.CalcDensityParticleAll(int(g["k"]))
    assert np.max(np.abs(rg / g["rho_gather"] - 1.0)) < tol
    lab, ng = t.FOF(float(g["fof_b"]), int(g["fof_minnum"]))
    assert ng == int(g["fof_ngroups"]) and np.array_equal(lab, g["fof_labels"])


@pytest.mark.gpu
def test_golden_phase6d_gpu(gpu):
    nb = _nb()
    g = _load("phase6d_n1024.npz")
    print("golden phase 6-D n=1024", flush=True)
    t = nb.KDTree(g["pos"], vel=g["vel"], treetype=nb.TPHS, kerntype=nb.KSPH, aniso=-1)
    ids, d2 = t.FindNearestPhase(int(g["k"]))
    assert _rows_equal(ids, g["ids"]) and np.array_equal(d2, g["d2"])

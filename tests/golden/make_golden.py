"""Writes the golden fixtures in this folder from the CPU oracle's brute-force
restatements (O(N^2) kNN with the reference's target/position rules and
sequential periodic images, density over those lists in long double, FOF by
brute-force union-find with canonical labels).  Inputs are stored with the
outputs, so the fixtures do not depend on the generators staying unchanged.

    python tests/golden/make_golden.py        (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from nbodylib_amd import datagen  # noqa: E402


def uniform_periodic():
    n, k = 1024, 16
    pos = datagen.uniform(n, 1.0, 7)
    fn_ids, fn_d2 = oracle.knn_brute(pos, k, mode=1, period=1.0)
    fp_ids, fp_d2 = oracle.knn_brute(pos, k, mode=2, period=1.0)
    lists, _ = oracle.knn_brute(pos, 32, mode=0)
    rho, h = oracle.density_from_lists(pos, None, lists)
    b = 0.5 * datagen.mean_spacing(n, 1.0)
    lab, ng = oracle.fof_brute(pos, b, 8, period=1.0)
    np.savez_compressed(os.path.join(HERE, "uniform_periodic_n1024.npz"), pos=pos, k=k, period=1.0,
                        findnearest_ids=fn_ids, findnearest_d2=fn_d2, findnearestpos_ids=fp_ids,
                        findnearestpos_d2=fp_d2, density_k=32, rho=rho, h=h, fof_b=b, fof_minnum=8, fof_labels=lab,
                        fof_ngroups=ng)


def clustered():
    n, k = 2048, 32
    pos, _ = datagen.clustered(n, 100.0, nhalo=20, seed=9)
    mass = datagen.lognormal_masses(n, seed=9)
    ids, _ = oracle.knn_brute(pos, k, mode=0)
    rho, h = oracle.density_from_lists(pos, mass, ids)
    rho_g, _ = oracle.density_from_lists(pos, mass, ids, symmetric=False)
    b = 0.2 * datagen.mean_spacing(n, 100.0)
    lab, ng = oracle.fof_brute(pos, b, 8)
    np.savez_compressed(os.path.join(HERE, "clustered_n2048.npz"), pos=pos, mass=mass, k=k, ids=ids, rho=rho, h=h,
                        rho_gather=rho_g, fof_b=b, fof_minnum=8, fof_labels=lab, fof_ngroups=ng)


def phase6d():
    n, k = 1024, 16
    pos, vel, _ = datagen.clustered(n, 100.0, nhalo=10, seed=13, with_vel=True)
    x = pos / pos.std(axis=0)
    v = vel / vel.std(axis=0)
    ids, d2 = oracle.knn_brute(x, k, mode=0, vel=v, nd=6)
    np.savez_compressed(os.path.join(HERE, "phase6d_n1024.npz"), pos=x, vel=v, k=k, ids=ids, d2=d2)


if __name__ == "__main__":
    oracle.build()
    uniform_periodic()
    clustered()
    phase6d()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))

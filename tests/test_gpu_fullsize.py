"""Size-independent properties at the headline size (16.8M clustered
particles, periodic L = 100), on the GPU:
  * FOF at b and FOFCriterion with FOF3d at params[6] = b^2 give the same
    labels (the same closure through two link metrics);
  * with unit velocities and rho = 1 given, CalcSmoothVel adds exactly
    CalcDensity's terms, so each component equals the density up to the
    order of the atomic sums;
  * on the 6-D phase-space set, FOFCriterion with FOF6d at scales (s6, sv)
    and plain FOF at b^2 = s6 on a tree whose velocities are scaled by
    sqrt(s6 / sv) give the same labels (the same closure through two link
    metrics and two trees).
Each case prints its name before the GPU call."""
import numpy as np
import pytest

from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu
N, L = 1 << 24, 100.0


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def clustered16m():
    pos, _ = datagen.clustered(N, L, nhalo=2000, seed=20261203)
    return pos


def test_fof_equals_fof3d_criterion(gpu, clustered16m):
    import nbodylib_amd as nb

    b = 0.2 * datagen.mean_spacing(N, L)
    say(f"fof vs FOF3d n={N} b={b:.5f}")
    t = nb.KDTree(clustered16m, kerntype=nb.KSPH, period=L)
    lab, ng = t.FOF(b, 8)
    lc, nc = t.FOFCriterion(nb.FOF_CMP_3D, [nb.TPHYS, b * b, 0, 0, 0, 0, b * b], 8)
    assert ng > 1000 and nc == ng
    assert np.array_equal(lab, lc)


def test_smooth_vel_unit_fields_is_density(gpu, clustered16m):
    import nbodylib_amd as nb

    k = 32
    say(f"smooth vel identity n={N} k={k}")
    t = nb.KDTree(clustered16m, vel=np.ones((N, 3)), kerntype=nb.KSPH, period=L)
    rho = t.CalcDensity(k)
    mean = t.CalcSmoothVel(k, rho=np.ones(N))
    assert np.max(np.abs(mean / rho[:, None] - 1.0)) < 1e-12


def test_fof6d_criterion_equals_scaled_plain_fof(gpu):
    import nbodylib_amd as nb

    pos, vel, _ = datagen.clustered(N, L, nhalo=2000, seed=20261203 + 2, with_vel=True)
    x = pos / pos.std(axis=0)
    v = vel / vel.std(axis=0)
    s6 = (0.2 * datagen.mean_spacing(N, L) / float(np.mean(pos.std(axis=0)))) ** 2
    sv = 0.01
    say(f"FOF6d vs scaled plain 6-D FOF n={N} s6={s6:.3e} sv={sv}")
    t = nb.KDTree(x, vel=v, treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    lc, nc = t.FOFCriterion(nb.FOF_CMP_6D, [nb.TPHS, s6, sv, 0, 0, 0, s6, sv], 8)
    del t
    ts = nb.KDTree(x, vel=v * np.sqrt(s6 / sv), treetype=nb.TPHS, aniso=-1, kerntype=nb.KSPH)
    lab, ng = ts.FOF(np.sqrt(s6), 8)
    assert nc > 1000 and ng == nc
    assert np.array_equal(lab, lc)

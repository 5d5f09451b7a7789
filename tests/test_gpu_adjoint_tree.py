"""Differentiation form over the tree's own neighbour lists recomputed per
block of tree slots (nbx_sph_density_fwd_tree_dev /
nbx_sph_density_adjoint_tree_dev) against the same forward and adjoint over
a stored list of all n (nbx_sph_density_fwd_dev / _adjoint_dev): densities,
smoothing lengths and gradients agree to 1e-12 (only the atomic summation
order differs), for several block sizes and both forms."""
import numpy as np
import pytest

from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_recompute_equals_stored(gpu, form):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import DeviceTree, _check, _ptr, _stream, lib

    n, k = 200000, 32
    say(f"adjoint recompute vs stored form={form} n={n} k={k}")
    pos = datagen.clustered(n, 100.0, nhalo=100, seed=71)[0]
    m = datagen.lognormal_masses(n, 71)
    pt = torch.tensor(np.ascontiguousarray(pos.T), dtype=torch.float64, device="cuda")
    mt = torch.tensor(m, dtype=torch.float64, device="cuda")
    f = nb.FORM_MIXED if form == "mixed" else nb.FORM_PLAIN
    t = DeviceTree(pt, mass=mt, kerntype=nb.KSPH, period=100.0, form=f)
    # the stored list: CalcDensity's search, rows by ID, ascending (FP64 kernel)
    nbr = torch.empty((n, k), dtype=torch.int32, device="cuda")
    d2 = torch.empty((n, k), dtype=torch.float64, device="cuda")
    t.knn(k, nbr, d2, variant=nb.KNN_NONPERIODIC, layout=nb.LAYOUT_BY_ID)
    rho0 = torch.empty(n, dtype=torch.float64, device="cuda")
    h0 = torch.empty(n, dtype=torch.float64, device="cuda")
    _check(lib().nbx_sph_density_fwd_dev(_ptr(pt), _ptr(mt), _ptr(nbr), n, k, _ptr(rho0), _ptr(h0), _stream()))
    g = 1.0 / rho0  # d(sum log rho)/d rho
    dp0 = torch.empty_like(pt)
    dm0 = torch.empty_like(mt)
    _check(lib().nbx_sph_density_adjoint_dev(_ptr(pt), _ptr(mt), _ptr(nbr), n, k, _ptr(g), _ptr(dp0), _ptr(dm0),
                                             _stream()))
    for block in (n, 65536, 5000):
        rho = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
        h = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
        t.sph_density(k, rho, h, block_rows=block)
        dp = torch.full_like(pt, 7.0)
        dm = torch.full_like(mt, 7.0)
        t.sph_density_adjoint(k, g, dp, dm, block_rows=block)
        torch.cuda.synchronize()
        say(f"  block={block}")
        assert torch.equal(h, h0)
        assert float(((rho - rho0).abs() / rho0).max()) < 1e-12
        for a, b in ((dp, dp0), (dm, dm0)):
            assert float((a - b).abs().max()) <= 1e-12 * float(b.abs().max())

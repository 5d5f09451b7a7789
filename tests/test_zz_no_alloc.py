"""nbx_knn_density_noalloc (the fused kNN + CalcDensity pass with its
temporaries in the library's file-scope device scratch) against
nbx_knn_density_dev, both through the C ABI: neighbour rows and smoothing
lengths identical, densities within 1e-12 relative (the scatter's atomic
order differs between runs), on two calls in a row, the second on a smaller
tree; and a tree too large for the scratch rejected before anything runs."""
import ctypes

import numpy as np
import pytest

from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu


def _run(entry, t, k):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import _ptr, _stream

    nn = torch.full((t.n, k), -7, dtype=torch.int32, device="cuda")
    rho = torch.full((t.n,), -1.0, dtype=torch.float64, device="cuda")
    h = torch.full((t.n,), -1.0, dtype=torch.float64, device="cuda")
    rc = getattr(nb.lib(), entry)(t._h, k, _ptr(nn), _ptr(rho), _ptr(h), _stream())
    torch.cuda.synchronize()
    return rc, nn.cpu().numpy(), rho.cpu().numpy(), h.cpu().numpy()


@pytest.mark.parametrize("form", ["mixed", "plain"])
def test_noalloc_matches_dev(gpu, form):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import DeviceTree

    f = nb.FORM_MIXED if form == "mixed" else nb.FORM_PLAIN
    for n, seed in ((200000, 61), (50000, 62)):
        print(f"noalloc form={form} n={n} k=32", flush=True)
        pos = datagen.clustered(n, 100.0, nhalo=max(4, n // 2000), seed=seed)[0]
        m = datagen.lognormal_masses(n, seed)
        pt = torch.tensor(np.ascontiguousarray(pos.T), dtype=torch.float64, device="cuda")
        mt = torch.tensor(m, dtype=torch.float64, device="cuda")
        t = DeviceTree(pt, mass=mt, kerntype=nb.KSPH, period=100.0, form=f)
        rc0, nn0, r0, h0 = _run("nbx_knn_density_dev", t, 32)
        rc1, nn1, r1, h1 = _run("nbx_knn_density_noalloc", t, 32)
        assert rc0 == 0 and rc1 == 0, nb.lib().nbx_last_error()
        assert np.array_equal(nn1, nn0)
        assert np.array_equal(h1, h0)
        assert np.all(r0 > 0) and np.max(np.abs(r1 / r0 - 1.0)) < 1e-12
        del t


def test_noalloc_rejects_oversize(gpu):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import DeviceTree

    n = 24_000_000  # 8 bytes each plus the precision form's 4-byte fallback list: 275 MiB > 256 MiB
    print(f"noalloc oversize n={n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(63)
    pt = torch.rand((3, n), dtype=torch.float64, device="cuda", generator=g)
    t = DeviceTree(pt, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    rc, nn, rho, h = _run("nbx_knn_density_noalloc", t, 32)
    assert rc == nb.NBX_ERR_UNSUPPORTED
    assert b"MiB" in nb.lib().nbx_last_error()
    assert np.all(nn == -7) and np.all(rho == -1.0)

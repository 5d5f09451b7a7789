"""Size-independent properties at the combined large size, N = 512^3 =
134,217,728 clustered particles (half uniform, half in 2000 NFW halos,
periodic L = 100, lognormal masses), on one GPU:
  * the precision form's neighbour rows (FP32 search with the FP64 re-check
    band) equal the plain FP64 form's on a sample of 1% of the particles,
    whose rows the FP64 kernel computes per target (nbx_knn_targets);
  * FOF at b = 0.2 x mean spacing and FOFCriterion with FOF3d at
    params[6] = b^2 give the same labels (the same closure through two link
    metrics);
  * the density with the list recomputed per block of tree slots equals the
    precision form's kNN + density fused pass's smoothing lengths exactly.
Each case prints its name before the GPU call."""
import numpy as np
import pytest

from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu
N, L, K = 512 ** 3, 100.0, 32


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c4(gpu):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import DeviceTree

    say(f"C4: generating clustered N={N}")
    pos = datagen.clustered(N, L, nhalo=2000, seed=20261207)[0]
    m = datagen.lognormal_masses(N, 20261207)
    pt = torch.from_numpy(np.ascontiguousarray(pos.T)).to("cuda")
    del pos
    mt = torch.from_numpy(m).to("cuda")
    del m
    say("C4: building the precision-form tree")
    t = DeviceTree(pt, mass=mt, bucket_size=16, kerntype=nb.KSPH, period=L, form=nb.FORM_MIXED)
    yield t
    del t


def test_c4_precision_rows_equal_fp64_sample(c4):
    import torch

    import nbodylib_amd as nb

    t = c4
    say(f"C4: fused kNN + density, precision form, N={N} k={K}")
    nn = torch.empty((N, K), dtype=torch.int32, device="cuda")
    rho = torch.empty(N, dtype=torch.float64, device="cuda")
    h = torch.empty(N, dtype=torch.float64, device="cuda")
    t.knn_density(K, nn, rho, h)
    torch.cuda.synchronize()
    order = t.order().cpu().numpy()  # slot -> ID
    slots = np.arange(0, N, 97, dtype=np.int64)
    ids = order[slots].astype(np.int32)
    rows = nn[torch.from_numpy(slots).to("cuda")].cpu().numpy()
    hs = h.cpu().numpy()[ids]
    del nn, rho
    say(f"C4: FP64 rows of {len(ids)} sampled targets")
    fi = np.zeros((len(ids), K), dtype=np.int32)
    fd = np.zeros((len(ids), K), dtype=np.float64)
    nb._check(nb.lib().nbx_knn_targets(t._h, nb._p(ids), len(ids), K, nb.KNN_NONPERIODIC, nb.LAYOUT_BY_SLOT,
                                       nb._p(fi), nb._p(fd)))
    assert np.array_equal(np.sort(rows, 1), np.sort(fi, 1))
    assert np.array_equal(hs, 0.5 * np.sqrt(fd[:, -1]))
    assert t.last_fallbacks() >= 0
    # the recomputed-list forward's h equals the search's h exactly
    say("C4: differentiation-form forward with the list recomputed per 4M-slot block")
    rho2 = torch.empty(N, dtype=torch.float64, device="cuda")
    h2 = torch.empty(N, dtype=torch.float64, device="cuda")
    t.sph_density(K, rho2, h2, block_rows=1 << 22)
    torch.cuda.synchronize()
    assert np.array_equal(h2.cpu().numpy()[ids], hs)
    assert float(rho2.min()) > 0


def test_c4_fof_equals_fof3d_criterion(c4):
    import torch

    import nbodylib_amd as nb

    t = c4
    b = 0.2 * datagen.mean_spacing(N, L)
    say(f"C4: FOF b={b} and FOFCriterion(FOF3d) params[6]=b^2, N={N}")
    l1 = torch.empty(N, dtype=torch.int32, device="cuda")
    l2 = torch.empty(N, dtype=torch.int32, device="cuda")
    g1 = t.fof(b, 8, l1)
    g2 = t.fof_criterion(nb.FOF_CMP_3D, [nb.TPHYS, b * b, 0, 0, 0, 0, b * b, 0], 8, l2)
    assert g1 == g2 and g1 > 1000
    assert torch.equal(l1, l2)

"""The device block a tree keeps between calls (densities by slot, fallback
list, search scratch, FOF work arrays): results must not depend on what an
earlier call left in it, on the block growing or being reused for a smaller
need, or on two threads calling on the same tree from their own streams."""
import threading

import numpy as np
import pytest

from nbodylib_amd import datagen

pytestmark = pytest.mark.gpu


def _tree(n, seed, form):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd.device import DeviceTree

    pos = datagen.clustered(n, 100.0, nhalo=max(4, n // 2000), seed=seed)[0]
    m = datagen.lognormal_masses(n, seed)
    pt = torch.tensor(np.ascontiguousarray(pos.T), dtype=torch.float64, device="cuda")
    mt = torch.tensor(m, dtype=torch.float64, device="cuda")
    f = nb.FORM_MIXED if form == "mixed" else nb.FORM_PLAIN
    return DeviceTree(pt, mass=mt, kerntype=nb.KSPH, period=100.0, form=f)


def _knn_density(t, k, stream=None):
    """The fused pass with its outputs filled, written and copied back all
    on one stream (the caller's, else the current one)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        nn = torch.full((t.n, k), -7, dtype=torch.int32, device="cuda")
        rho = torch.full((t.n,), -1.0, dtype=torch.float64, device="cuda")
        h = torch.full((t.n,), -1.0, dtype=torch.float64, device="cuda")
        t.knn_density(k, nn, rho, h)
        out = nn.cpu().numpy(), rho.cpu().numpy(), h.cpu().numpy()
    s.synchronize()
    return out


def _fof(t, b, minnum=8):
    import torch

    lab = torch.full((t.n,), -3, dtype=torch.int32, device="cuda")
    ng = t.fof(b, minnum, lab)
    torch.cuda.synchronize()
    return ng, lab.cpu().numpy()


def _same(a, b):
    nn0, r0, h0 = a
    nn1, r1, h1 = b
    assert np.array_equal(nn1, nn0)
    assert np.array_equal(h1, h0)
    # the density scatter's atomic order differs from run to run
    assert np.all(r0 > 0) and np.max(np.abs(r1 / r0 - 1.0)) < 1e-12


@pytest.mark.parametrize("form", ["mixed", "plain"])
def test_kept_scratch_sequence(gpu, form):
    """Density (first allocation), FOF (block grows for FOF's arrays), a
    density with larger k (precision form: larger search scratch), then the
    first call again on a block now larger than it needs."""
    n = 120000
    print(f"kept scratch sequence form={form} n={n}", flush=True)
    t = _tree(n, 71, form)
    b = 0.2 * datagen.mean_spacing(n, 100.0)
    first = _knn_density(t, 32)
    ng0, lab0 = _fof(t, b)
    big = _knn_density(t, 64)
    again = _knn_density(t, 32)
    _same(first, again)
    ng1, lab1 = _fof(t, b)
    assert ng1 == ng0 and ng0 > 0
    assert np.array_equal(lab1, lab0)
    # the k=64 rows start with the k=32 rows' sets (both ascending by (d2, ID) in the FP64 form's order)
    assert big[0].shape == (n, 64)
    assert np.all(big[2] >= first[2])


def test_kept_scratch_two_threads(gpu):
    """Two threads call the fused pass on one tree, each on its own stream;
    the calls take turns with the block and both give the serial answer."""
    import torch

    n = 150000
    print(f"kept scratch two threads n={n}", flush=True)
    t = _tree(n, 72, "mixed")
    ref = _knn_density(t, 32)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    out = [None, None]
    errs = []

    def work(i):
        try:
            torch.cuda.set_device(0)
            for _ in range(3):
                out[i] = _knn_density(t, 32, streams[i])
        except Exception as e:  # reported below, on the main thread
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    assert not any(x.is_alive() for x in th), "a caller did not return"
    assert not errs, errs
    for o in out:
        _same(ref, o)

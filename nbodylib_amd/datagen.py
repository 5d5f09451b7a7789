"""Seeded synthetic particle sets for the KDTree hot path.

* uniform(n, L, seed): positions U[0, L)^3.
* clustered(n, L, seed): half uniform background, half in 2000 NFW halos
  (concentration U(5,15), scale radius U(0.05,0.5)*L/100, halo masses from
  dN/dM ~ M^-1.9 over three decades, member counts proportional to mass by
  largest-remainder rounding, radii from the NFW enclosed-mass profile
  mu(x) = ln(1+x) - x/(1+x) inverted by bisection and truncated at c,
  isotropic directions), wrapped into the periodic box.
* velocities: halo members get the halo's bulk velocity N(0,300)^3 plus an
  isotropic dispersion sigma_h = 300 (M_h/M_max)^(1/3) (V_vir ~ M^(1/3));
  background particles N(0,300)^3.
* lognormal_masses: m_i = exp(0.3 g - 0.045)/n (mean 1/n).

All draws come from numpy's PCG64 generator, so a seed fixes the set.
"""
import numpy as np


def _wrap(x, L):
    x = np.mod(x, L)
    x[x >= L] = 0.0
    return x


def uniform(n, L=1.0, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)) * L


def _nfw_x_chunk(u, c, iters):
    mu = lambda x: np.log1p(x) - x / (1.0 + x)
    target = u * mu(c)
    lo = np.zeros_like(target)
    hi = np.array(c, dtype=np.float64) * np.ones_like(target)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        f = mu(mid) - target
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return 0.5 * (lo + hi)


def _nfw_x(u, c, iters=60):
    """Solve mu(x) = u * mu(c) for x in (0, c] by bisection (vectorised,
    elementwise, so chunks run on a thread pool with identical results)."""
    from concurrent.futures import ThreadPoolExecutor
    import os

    n = len(u)
    step = 1 << 20
    if n <= step:
        return _nfw_x_chunk(u, c, iters)
    out = np.empty(n, dtype=np.float64)

    def work(a):
        out[a:a + step] = _nfw_x_chunk(u[a:a + step], c[a:a + step], iters)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(work, range(0, n, step)))
    return out


def clustered(n, L=100.0, nhalo=2000, seed=0, with_vel=False, alpha=1.9, mrange=1e3):
    """Returns pos (n,3) [and vel (n,3), halo id per particle (-1 background)]."""
    rng = np.random.default_rng(seed)
    nbg = n // 2
    nh_tot = n - nbg
    pos = np.empty((n, 3), dtype=np.float64)
    pos[:nbg] = rng.random((nbg, 3)) * L
    centres = rng.random((nhalo, 3)) * L
    conc = rng.uniform(5.0, 15.0, nhalo)
    rs = rng.uniform(0.05, 0.5, nhalo) * (L / 100.0)
    # masses from dN/dM ~ M^-alpha on [1, mrange] (inverse CDF)
    a1 = 1.0 - alpha
    u = rng.random(nhalo)
    M = (1.0 + u * (mrange ** a1 - 1.0)) ** (1.0 / a1)
    share = M / M.sum() * nh_tot
    cnt = np.floor(share).astype(np.int64)
    rem = nh_tot - int(cnt.sum())
    if rem > 0:
        order = np.argsort(-(share - cnt), kind="stable")
        cnt[order[:rem]] += 1
    hid = np.repeat(np.arange(nhalo), cnt)
    x = _nfw_x(rng.random(nh_tot), conc[hid])
    r = x * rs[hid]
    v = rng.normal(size=(nh_tot, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pos[nbg:] = centres[hid] + r[:, None] * v
    pos = _wrap(pos, L)
    halo = np.concatenate([np.full(nbg, -1, dtype=np.int64), hid])
    if not with_vel:
        return pos, halo
    vel = np.empty((n, 3), dtype=np.float64)
    vel[:nbg] = rng.normal(0.0, 300.0, (nbg, 3))
    bulk = rng.normal(0.0, 300.0, (nhalo, 3))
    sig = 300.0 * (M / M.max()) ** (1.0 / 3.0)
    vel[nbg:] = bulk[hid] + rng.normal(size=(nh_tot, 3)) * sig[hid][:, None]
    return pos, vel, halo


def lognormal_masses(n, seed=0, sigma=0.3):
    rng = np.random.default_rng(seed + 7919)
    return np.exp(sigma * rng.normal(size=n) - 0.5 * sigma * sigma) / n


def mean_spacing(n, L):
    return L / n ** (1.0 / 3.0)

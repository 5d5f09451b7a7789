"""ctypes binding of the CPU oracle (oracle/nbx_oracle.c).

Test infrastructure only: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py.  The shipped library never imports it.

Arrays are numpy: positions are passed as an (n, 6) float64 "phase" array
(x y z vx vy vz, row-major) the way the reference's Particle stores them
(src/NBody/Particle.h:264-293); outputs are keyed by input index (= the ID
the reference writes in KDTree.cxx:1291).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_ci = ctypes.c_int
_cd = ctypes.c_double

TPHYS, TPHS = 0, 3
KSPH, KGAUSS, KEPAN, KTH = 0, 1, 2, 3


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_set_mutation.argtypes = [_ci]
        L.orc_kernel.argtypes = [_ci, _ci, _ci, ctypes.POINTER(_cd), ctypes.c_void_p]
        L.orc_kernel.restype = _ci
        L.orc_tree_build.argtypes = [_dp, ctypes.c_void_p, _i64, _ci, _ci, _ci, _ci, ctypes.c_void_p]
        L.orc_tree_build.restype = _vp
        L.orc_tree_free.argtypes = [_vp]
        L.orc_tree_numnodes.argtypes = [_vp]
        L.orc_tree_numnodes.restype = _i64
        L.orc_tree_numleaves.argtypes = [_vp]
        L.orc_tree_numleaves.restype = _i64
        L.orc_tree_order.argtypes = [_vp, _ip]
        L.orc_knn_tree.argtypes = [_vp, _ci, _ci, _ip, ctypes.c_void_p, _ci]
        L.orc_knn_points.argtypes = [_vp, _dp, _i64, _ci, _ip, ctypes.c_void_p, _ci]
        L.orc_calc_density.argtypes = [_vp, _ci, _dp]
        L.orc_calc_density.restype = _ci
        L.orc_calc_density_gather.argtypes = [_vp, _ci, _dp, ctypes.c_void_p, _ci]
        L.orc_calc_density_gather.restype = _ci
        L.orc_calc_veldensity.argtypes = [_vp, _ci, _ci, _dp, _ci]
        L.orc_calc_veldensity.restype = _ci
        L.orc_calc_veldensity_position.argtypes = [_vp, _dp, _dp, ctypes.c_int64, _ci, _ci, _dp, _vp, _ci]
        L.orc_calc_veldensity_position.restype = _ci
        L.orc_calc_veldensity_phys.argtypes = [_vp, _ci, _ci, _dp, _dp, _ci]
        L.orc_calc_veldensity_phys.restype = _ci
        L.orc_calc_density_position.argtypes = [_vp, _dp, _i64, _ci, _dp, ctypes.c_void_p, _ci]
        L.orc_calc_density_position.restype = _ci
        L.orc_smooth_vel.argtypes = [_vp, _ci, ctypes.c_void_p, _dp, ctypes.c_void_p]
        L.orc_smooth_vel.restype = _ci
        L.orc_smooth_moments.argtypes = [_vp, _ci, ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_smooth_moments.restype = _ci
        L.orc_smooth_local.argtypes = [_vp, ctypes.c_void_p, _i64, _ci, _dp, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        L.orc_smooth_local.restype = _ci
        L.orc_fof_tree.argtypes = [_vp, _cd, _ci, _ci, _ip]
        L.orc_fof_tree.restype = _i64
        L.orc_knn_brute.argtypes = [_dp, _i64, _ci, _ci, _ci, ctypes.c_void_p, _ip, ctypes.c_void_p, _ci]
        L.orc_density_from_lists.argtypes = [_dp, ctypes.c_void_p, _i64, _ci, _ci, _ip, _dp, _ci, _ci, _dp,
                                             ctypes.c_void_p]
        L.orc_fof_canonical.argtypes = [_i64, _ip, _ci, _ip]
        L.orc_fof_canonical.restype = _i64
        L.orc_fof_brute.argtypes = [_dp, _i64, _cd, _ci, ctypes.c_void_p, _ip]
        L.orc_fof_brute.restype = _i64
        L.orc_fof_brute_nd.argtypes = [_dp, _i64, _ci, _cd, _ci, ctypes.c_void_p, _ip]
        L.orc_fof_brute_nd.restype = _i64
        L.orc_fof_criterion_brute.argtypes = [_dp, _i64, _ci, _dp, _ci, ctypes.c_void_p, _ip]
        L.orc_fof_criterion_brute.restype = _i64
        L.orc_ball_mark.argtypes = [_dp, _i64, _dp, _i64, _cd, ctypes.c_void_p, _ci, _ip, _dp]
        L.orc_ball_mark.restype = None
        L.orc_density_analytic.argtypes = [_dp, _dp, _i64, _ip, _ci, _dp, ctypes.c_void_p]
        L.orc_density_adjoint.argtypes = [_dp, _dp, _i64, _ip, _ci, _dp, _dp, _dp]
        L.orc_knn_density_sample.argtypes = [_vp, _ci, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                             _i64, _dp, _ci]
        L.orc_knn_density_sample.restype = _i64
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = _ci
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _phase(pos, vel=None):
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    ph = np.zeros((n, 6), dtype=np.float64)
    ph[:, :3] = pos
    if vel is not None:
        ph[:, 3:] = np.asarray(vel, dtype=np.float64)
    return np.ascontiguousarray(ph)


def set_mutation(m):
    lib().orc_set_mutation(int(m))


def kernel(nd, kerntype=KSPH, kernres=1000):
    """KernelConstruction (KDTree.cxx:1144-1183): returns (kernnorm, table)."""
    kn = ctypes.c_double()
    res = lib().orc_kernel(nd, kerntype, kernres, ctypes.byref(kn), None)
    tab = np.zeros(res, dtype=np.float64)
    lib().orc_kernel(nd, kerntype, res, None, _ptr(tab))
    return kn.value, tab


class Tree:
    """Faithful restatement of NBody::KDTree (median split, b-bucket leaves)."""

    def __init__(self, pos, mass=None, vel=None, b=16, treetype=TPHYS, kerntype=KSPH, kernres=1000, period=None):
        self.ph = _phase(pos, vel)
        self.n = self.ph.shape[0]
        self.mass = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self.period = None if period is None else np.ascontiguousarray(np.broadcast_to(period, (3,)), dtype=np.float64)
        self.h = lib().orc_tree_build(self.ph, _ptr(self.mass), self.n, b, treetype, kerntype, kernres,
                                      _ptr(self.period))
        if not self.h:
            raise ValueError("invalid tree type")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    @property
    def numnodes(self):
        return lib().orc_tree_numnodes(self.h)

    @property
    def numleaves(self):
        return lib().orc_tree_numleaves(self.h)

    def order(self):
        o = np.zeros(self.n, dtype=np.int32)
        lib().orc_tree_order(self.h, o)
        return o

    def knn(self, k, variant=0, nthreads=0):
        """variant 0 FindNearest(tt), 1 FindNearestPos(tt), 2 FindNearestPhase(tt),
        3 the non-periodic search CalcDensity uses.  Rows keyed by ID."""
        ids = np.zeros((self.n, k), dtype=np.int32)
        d2 = np.zeros((self.n, k), dtype=np.float64)
        lib().orc_knn_tree(self.h, k, variant, ids, _ptr(d2), nthreads)
        return ids, d2

    def knn_points(self, q, k, nthreads=0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        m = q.shape[0]
        ids = np.zeros((m, k), dtype=np.int32)
        d2 = np.zeros((m, k), dtype=np.float64)
        lib().orc_knn_points(self.h, q, m, k, ids, _ptr(d2), nthreads)
        return ids, d2

    def calc_density(self, k):
        rho = np.zeros(self.n, dtype=np.float64)
        rc = lib().orc_calc_density(self.h, k, rho)
        if rc != 0:
            raise RuntimeError("CalcDensity failed")
        return rho

    def calc_density_gather(self, k, nthreads=0):
        rho = np.zeros(self.n, dtype=np.float64)
        h = np.zeros(self.n, dtype=np.float64)
        rc = lib().orc_calc_density_gather(self.h, k, rho, _ptr(h), nthreads)
        if rc != 0:
            raise RuntimeError("CalcDensityParticle failed")
        return rho, h

    def calc_density_position(self, q, k, nthreads=0):
        """CalcDensityPosition(x, k) at each row of q (m, 3): (rho, h)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        m = q.shape[0]
        rho = np.zeros(m, dtype=np.float64)
        h = np.zeros(m, dtype=np.float64)
        rc = lib().orc_calc_density_position(self.h, q, m, k, rho, _ptr(h), nthreads)
        if rc != 0:
            raise RuntimeError("CalcDensityPosition failed")
        return rho, h

    def smooth_vel(self, k, rho=None, disp=False):
        """CalcSmoothVel(k) [and CalcSmoothVelDisp]: mean (n, 3) [, disp (n, 3, 3)] by ID; rho by ID or None
        (CalcDensity(k) first)."""
        mean = np.zeros((self.n, 3), dtype=np.float64)
        dd = np.zeros((self.n, 9), dtype=np.float64) if disp else None
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        rc = lib().orc_smooth_vel(self.h, k, _ptr(r), mean, _ptr(dd))
        if rc != 0:
            raise RuntimeError("CalcSmoothVel failed")
        return (mean, dd.reshape(self.n, 3, 3)) if disp else mean

    def smooth_moments(self, k, mean, disp, rho=None):
        """CalcSmoothVelSkew / CalcSmoothVelKurtosis given means (n, 3) and dispersions (n, 3, 3) by ID:
        (skew, kurt), each (n, 3)."""
        skew = np.zeros((self.n, 3), dtype=np.float64)
        kurt = np.zeros((self.n, 3), dtype=np.float64)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        mn = np.ascontiguousarray(mean, dtype=np.float64).reshape(self.n, 3)
        dp = np.ascontiguousarray(disp, dtype=np.float64).reshape(self.n, 9)
        rc = lib().orc_smooth_moments(self.h, k, _ptr(r), mn, dp, skew, kurt)
        if rc != 0:
            raise RuntimeError("CalcSmoothVelSkew failed")
        return skew, kurt

    def smooth_local(self, k, weight, rho=None, mean_in=None, q=None):
        """CalcSmoothLocalMean/DispParticle for every particle (q None; by ID) or the Position forms at the rows
        of q (m, 3): (mean, disp); disp only when mean_in is given."""
        w = np.ascontiguousarray(weight, dtype=np.float64)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        rows = self.n if qq is None else qq.shape[0]
        mi = None if mean_in is None else np.ascontiguousarray(mean_in, dtype=np.float64)
        mean = np.zeros(rows)
        disp = np.zeros(rows) if mi is not None else None
        rc = lib().orc_smooth_local(self.h, _ptr(qq), rows, k, w, _ptr(r), _ptr(mi), _ptr(mean), _ptr(disp))
        if rc != 0:
            raise RuntimeError("CalcSmoothLocalMean failed")
        return mean, disp

    def calc_veldensity(self, nsmooth, nsearch, nthreads=0):
        rho = np.zeros(self.n, dtype=np.float64)
        rc = lib().orc_calc_veldensity(self.h, nsmooth, nsearch, rho, nthreads)
        if rc != 0:
            raise RuntimeError("CalcVelDensity failed")
        return rho

    def calc_veldensity_phys(self, nsmooth, nsearch, rho_by_id, nthreads=0):
        """CalcVelDensityWithPhysDensity(Nsmooth, Nsearch) over the given densities; keyed by ID."""
        out = np.zeros(self.n, dtype=np.float64)
        rc = lib().orc_calc_veldensity_phys(self.h, nsmooth, nsearch,
                                            np.ascontiguousarray(rho_by_id, dtype=np.float64), out, nthreads)
        if rc != 0:
            raise RuntimeError("CalcVelDensityWithPhysDensity failed")
        return out

    def calc_veldensity_position(self, q, qv, nsmooth, nsearch, nthreads=0):
        """CalcVelDensityPosition(x, v, Nsmooth, Nsearch) at each row of q, qv (m, 3): (rho, h_v)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        qv = np.ascontiguousarray(qv, dtype=np.float64)
        m = q.shape[0]
        rho = np.zeros(m, dtype=np.float64)
        h = np.zeros(m, dtype=np.float64)
        rc = lib().orc_calc_veldensity_position(self.h, q, qv, m, nsmooth, nsearch, rho, _ptr(h), nthreads)
        if rc != 0:
            raise RuntimeError("CalcVelDensityPosition failed")
        return rho, h

    def knn_density_sample(self, k, slots, nthreads=0):
        """kNN + gather density for the given tree slots (CPU baseline sample)."""
        slots = np.ascontiguousarray(slots, dtype=np.int64)
        rho = np.zeros(slots.shape[0], dtype=np.float64)
        lib().orc_knn_density_sample(self.h, k, slots, slots.shape[0], rho, nthreads)
        return rho

    def fof(self, b, minnum=8, order=0):
        lab = np.zeros(self.n, dtype=np.int32)
        ng = lib().orc_fof_tree(self.h, b, minnum, order, lab)
        return lab, ng


def num_threads():
    return lib().orc_num_threads()


def knn_brute(pos, k, mode=0, period=None, vel=None, nd=3, nthreads=0):
    """mode 0: target search (no self, no d2==0); 1: periodic FindNearest;
    2: periodic FindNearestPos.  nd=6 uses the phase metric."""
    ph = _phase(pos, vel)
    n = ph.shape[0]
    ids = np.zeros((n, k), dtype=np.int32)
    d2 = np.zeros((n, k), dtype=np.float64)
    per = None if period is None else np.ascontiguousarray(np.broadcast_to(period, (3,)), dtype=np.float64)
    lib().orc_knn_brute(ph, n, nd, k, mode, _ptr(per), ids, _ptr(d2), nthreads)
    return ids, d2


def density_from_lists(pos, mass, ids, nd=3, kerntype=KSPH, kernres=1000, symmetric=True, vel=None):
    ph = _phase(pos, vel)
    n, k = ids.shape
    _, tab = kernel(nd, kerntype, kernres)
    rho = np.zeros(n, dtype=np.float64)
    h = np.zeros(n, dtype=np.float64)
    m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    lib().orc_density_from_lists(ph, _ptr(m), n, nd, k, np.ascontiguousarray(ids, dtype=np.int32), tab, len(tab),
                                 1 if symmetric else 0, rho, _ptr(h))
    return rho, h


def fof_canonical(labels, minnum=1):
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    out = np.zeros_like(labels)
    g = lib().orc_fof_canonical(labels.shape[0], labels, minnum, out)
    return out, g


def fof_brute(pos, b, minnum=8, period=None, vel=None):
    """All-pairs FOF, canonical labels.  With vel, links in 6-D phase space
    with the reference's d2 = |dx|^2 + |dv|^2 (periodic images shift x only)."""
    ph = _phase(pos, vel)
    n = ph.shape[0]
    lab = np.zeros(n, dtype=np.int32)
    per = None if period is None else np.ascontiguousarray(np.broadcast_to(period, (3,)), dtype=np.float64)
    g = lib().orc_fof_brute_nd(ph, n, 6 if vel is not None else 3, b, minnum, _ptr(per), lab)
    return lab, g


def fof_criterion_brute(pos, cmp, params, minnum=8, period=None, vel=None):
    """All-pairs FOFCriterion with FOF3d (cmp 1) or FOF6d (cmp 2, needs vel); canonical labels."""
    ph = _phase(pos, vel)
    n = ph.shape[0]
    p = np.zeros(8, dtype=np.float64)
    v = np.asarray(params, dtype=np.float64).ravel()[:8]
    p[: v.size] = v
    lab = np.zeros(n, dtype=np.int32)
    per = None if period is None else np.ascontiguousarray(np.broadcast_to(period, (3,)), dtype=np.float64)
    g = lib().orc_fof_criterion_brute(ph, n, cmp, p, minnum, _ptr(per), lab)
    return lab, g


def ball_mark(pos, x, r2, imark, group, d2, exclude=-1, period=None):
    """SearchBallPos restated over all particles: group[id] = imark, d2[id] = d2 (in place, by ID)."""
    ph = _phase(pos)
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))
    per = None if period is None else np.ascontiguousarray(np.broadcast_to(period, (3,)), dtype=np.float64)
    lib().orc_ball_mark(ph, ph.shape[0], x, int(exclude), float(r2), _ptr(per), int(imark), group, d2)


def density_analytic(pos, mass, nbr):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    n, k = nbr.shape
    rho = np.zeros(n)
    h = np.zeros(n)
    lib().orc_density_analytic(pos, mass, n, nbr, k, rho, _ptr(h))
    return rho, h


def density_adjoint(pos, mass, nbr, g):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    g = np.ascontiguousarray(g, dtype=np.float64)
    n, k = nbr.shape
    dpos = np.zeros((n, 3))
    dm = np.zeros(n)
    lib().orc_density_adjoint(pos, mass, n, nbr, k, g, dpos, dm)
    return dpos, dm

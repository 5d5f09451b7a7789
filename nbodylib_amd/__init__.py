"""nbodylib_amd -- MI355X-native KDTree hot path with NBodylib's call surface.

Python mirror of NBody::KDTree (src/KDTree/KDTree.h:81-657) over the C ABI in
include/nbx.h (libnbx.so, hand-written HIP for gfx950).  Method names and
argument meanings follow the reference; results are keyed by particle ID
(= input index, KDTree.cxx:1291) unless a method says otherwise.

There is no CPU fallback: every call runs the HIP kernels, and importing the
module fails if libnbx.so has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libnbx.so")

# KDTree.h:90-97 constants
TPHYS, TPROJ, TVEL, TPHS, TMETRIC = 0, 1, 2, 3, 4
KSPH, KGAUSS, KEPAN, KTH = 0, 1, 2, 3
FORM_PLAIN, FORM_MIXED = 0, 1
KNN_FINDNEAREST, KNN_FINDNEARESTPOS, KNN_FINDNEARESTPHASE, KNN_NONPERIODIC = 0, 1, 2, 3
LAYOUT_BY_ID, LAYOUT_BY_SLOT = 0, 1
DENS_SYMMETRIC, DENS_GATHER, DENS_VEL = 0, 1, 2
FOF_CMP_3D, FOF_CMP_6D = 1, 2  # the built-in FOFCriterion comparators FOF3d, FOF6d

NBX_OK = 0
NBX_ERR_ARG, NBX_ERR_HIP, NBX_ERR_UNSUPPORTED, NBX_ERR_NOMEM, NBX_ERR_SEARCH = -1, -2, -3, -4, -5

EXPORTS = [
    "nbx_params_default", "nbx_tree_create", "nbx_tree_create_dev", "nbx_tree_destroy", "nbx_tree_order",
    "nbx_tree_size", "nbx_tree_kernel", "nbx_knn", "nbx_knn_dev", "nbx_knn_points", "nbx_knn_targets", "nbx_knn_phase_points", "nbx_density_points", "nbx_veldensity_points", "nbx_veldensity_phys", "nbx_smooth_vel", "nbx_smooth_vel_moments", "nbx_smooth_local", "nbx_smooth_local_points", "nbx_density",
    "nbx_density_dev", "nbx_knn_density_dev", "nbx_knn_density_noalloc", "nbx_ball_points", "nbx_ball_points_many", "nbx_ball_mark", "nbx_fof", "nbx_fof_dev", "nbx_fof_criterion", "nbx_fof_criterion_dev", "nbx_fof_lists",
    "nbx_sph_density_fwd_dev", "nbx_sph_density_adjoint_dev", "nbx_sph_density_fwd_tree_dev",
    "nbx_sph_density_adjoint_tree_dev", "nbx_prof_enable", "nbx_prof_reset", "nbx_prof_get",
    "nbx_tree_last_fallbacks", "nbx_last_error",
]


class NbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nbx error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """nbx_params (include/nbx.h), mirroring the KDTree constructor options."""
    _fields_ = [
        ("bucket_size", ctypes.c_int32), ("treetype", ctypes.c_int32), ("kerntype", ctypes.c_int32),
        ("kernres", ctypes.c_int32), ("splitting", ctypes.c_int32), ("aniso", ctypes.c_int32),
        ("scalespace", ctypes.c_int32), ("periodic", ctypes.c_int32), ("period", ctypes.c_double * 3),
        ("form", ctypes.c_int32), ("recheck_rel", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libnbx.so (raises if it is missing -- no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIBPATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    L.nbx_params_default.argtypes = [P(Params)]
    L.nbx_params_default.restype = None
    L.nbx_tree_create.argtypes = [P(Params), vp, vp, vp, i64, P(vp)]
    L.nbx_tree_create_dev.argtypes = [P(Params), vp, vp, vp, i64, vp, P(vp)]
    L.nbx_tree_destroy.argtypes = [vp]
    L.nbx_tree_order.argtypes = [vp, vp]
    L.nbx_tree_size.argtypes = [vp]
    L.nbx_tree_size.restype = i64
    L.nbx_tree_kernel.argtypes = [vp, P(dbl), vp, i32]
    L.nbx_knn.argtypes = [vp, i32, i32, i32, vp, vp]
    L.nbx_knn_dev.argtypes = [vp, i32, i32, i32, vp, vp, vp]
    L.nbx_knn_points.argtypes = [vp, vp, i64, i32, vp, vp]
    L.nbx_knn_phase_points.argtypes = [vp, vp, vp, i64, i32, vp, vp]
    L.nbx_knn_targets.argtypes = [vp, vp, i64, i32, i32, i32, vp, vp]
    L.nbx_density_points.argtypes = [vp, vp, i64, i32, vp, vp]
    L.nbx_veldensity_points.argtypes = [vp, vp, vp, i64, i32, i32, vp, vp]
    L.nbx_veldensity_phys.argtypes = [vp, i32, i32, vp, vp]
    L.nbx_smooth_vel.argtypes = [vp, i32, vp, vp, vp, vp]
    L.nbx_smooth_vel_moments.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.nbx_smooth_local.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.nbx_smooth_local_points.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, vp]
    L.nbx_density.argtypes = [vp, i32, i32, i32, vp, vp]
    L.nbx_density_dev.argtypes = [vp, i32, i32, i32, vp, vp, vp]
    L.nbx_knn_density_dev.argtypes = [vp, i32, vp, vp, vp, vp]
    L.nbx_knn_density_noalloc.argtypes = [vp, i32, vp, vp, vp, vp]
    L.nbx_ball_points.argtypes = [vp, vp, dbl, vp, i64, P(i64)]
    L.nbx_ball_mark.argtypes = [vp, vp, i64, dbl, i32, vp, vp]
    L.nbx_ball_points_many.argtypes = [vp, vp, i64, dbl, vp, vp, vp, i64]
    L.nbx_fof.argtypes = [vp, dbl, i32, vp, P(i64)]
    L.nbx_fof_dev.argtypes = [vp, dbl, i32, vp, P(i64), vp]
    L.nbx_fof_criterion.argtypes = [vp, i32, vp, i32, vp, P(i64)]
    L.nbx_fof_lists.argtypes = [vp, vp, i64, i64, vp, vp, vp, vp]
    L.nbx_fof_criterion_dev.argtypes = [vp, i32, vp, i32, vp, P(i64), vp]
    L.nbx_sph_density_fwd_dev.argtypes = [vp, vp, vp, i64, i32, vp, vp, vp]
    L.nbx_sph_density_adjoint_dev.argtypes = [vp, vp, vp, i64, i32, vp, vp, vp, vp]
    L.nbx_sph_density_fwd_tree_dev.argtypes = [vp, i32, i64, vp, vp, vp]
    L.nbx_sph_density_adjoint_tree_dev.argtypes = [vp, i32, i64, vp, vp, vp, vp]
    L.nbx_prof_enable.argtypes = [ctypes.c_int]
    L.nbx_prof_enable.restype = None
    L.nbx_prof_reset.argtypes = []
    L.nbx_prof_reset.restype = None
    L.nbx_prof_get.argtypes = [ctypes.c_char_p, P(dbl), P(i64)]
    L.nbx_tree_last_fallbacks.argtypes = [vp, P(i64)]
    L.nbx_last_error.argtypes = []
    L.nbx_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise NbxError(rc, lib().nbx_last_error().decode())
    return rc


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _soa(a, n, name):
    a = np.asarray(a, dtype=np.float64)
    if a.shape == (n, 3):
        a = a.T
    if a.shape != (3, n):
        raise ValueError(f"{name} must have shape (n, 3) or (3, n)")
    return np.ascontiguousarray(a)


def params(bucket_size=16, treetype=TPHYS, kerntype=KEPAN, kernres=1000, splitting=0, aniso=0, scalespace=0,
           period=None, form=FORM_PLAIN, recheck_rel=1e-5):
    """Build nbx_params with the reference constructor's defaults (KDTree.h:229-245)."""
    p = Params()
    lib().nbx_params_default(ctypes.byref(p))
    p.bucket_size, p.treetype, p.kerntype, p.kernres = bucket_size, treetype, kerntype, kernres
    p.splitting, p.aniso, p.scalespace = splitting, aniso, scalespace
    if period is not None:
        per = np.broadcast_to(np.asarray(period, dtype=np.float64), (3,))
        p.periodic = 1
        for i in range(3):
            p.period[i] = float(per[i])
    p.form, p.recheck_rel = form, recheck_rel
    return p


class KDTree:
    """NBody::KDTree over host numpy arrays (the constructor of KDTree.h:229-245).

    pos, vel: (n, 3) or (3, n) float64; mass: (n,) or None (equal masses 1/n).
    treetype=TVEL builds the tree on vel (pos may be None); its searches and
    densities then run in velocity space, as the reference's TVEL tree does.
    """

    def __init__(self, pos, mass=None, vel=None, bucket_size=16, treetype=TPHYS, kerntype=KEPAN, kernres=1000,
                 splitting=0, aniso=0, scalespace=0, period=None, form=FORM_PLAIN, recheck_rel=1e-5):
        src = vel if treetype == TVEL and vel is not None else pos
        if src is None:
            raise ValueError("pos is required (vel for a TVEL tree)")
        src = np.asarray(src, dtype=np.float64)
        n = src.shape[0] if src.ndim == 2 and src.shape[1] == 3 else src.shape[-1]
        self.n = int(n)
        self.treetype = treetype
        self.period = period
        self._params = params(bucket_size, treetype, kerntype, kernres, splitting, aniso, scalespace, period, form,
                              recheck_rel)
        spos = None if pos is None or treetype == TVEL else _soa(pos, n, "pos")
        svel = None if vel is None else _soa(vel, n, "vel")
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(lib().nbx_tree_create(ctypes.byref(self._params), _p(spos), _p(svel), _p(m), self.n, ctypes.byref(h)))
        self._h = h
        self._destroy = lib().nbx_tree_destroy

    def __del__(self):
        # the destroy function is bound at creation: at interpreter exit the
        # module's globals may already be gone
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self._h = None

    # -------------------------------------------------------------- info
    def order(self):
        """slot -> ID (the tree order the shim imposes on Particle arrays)."""
        o = np.zeros(self.n, dtype=np.int32)
        _check(lib().nbx_tree_order(self._h, _p(o)))
        return o

    def GetKernNorm(self):
        kn = ctypes.c_double()
        _check(lib().nbx_tree_kernel(self._h, ctypes.byref(kn), None, 0))
        return kn.value

    def kernel_table(self):
        kn = ctypes.c_double()
        res = _check(lib().nbx_tree_kernel(self._h, ctypes.byref(kn), None, 0))
        tab = np.zeros(res, dtype=np.float64)
        lib().nbx_tree_kernel(self._h, None, _p(tab), res)
        return tab

    def last_fallbacks(self):
        """Precision form: queries of the last call re-run in FP64 (certificate failed)."""
        v = ctypes.c_int64()
        _check(lib().nbx_tree_last_fallbacks(self._h, ctypes.byref(v)))
        return v.value

    # -------------------------------------------------------------- kNN
    def _knn(self, k, variant, layout=LAYOUT_BY_ID, with_d2=True):
        nn = np.zeros((self.n, k), dtype=np.int32)
        d2 = np.zeros((self.n, k), dtype=np.float64) if with_d2 else None
        _check(lib().nbx_knn(self._h, k, variant, layout, _p(nn), _p(d2)))
        return nn, d2

    def FindNearest(self, Nsearch=64, layout=LAYOUT_BY_ID):
        """FindNearest(Int_t**,Double_t**,Nsearch) for every particle (KDFindNearest.cxx:444-446)."""
        return self._knn(Nsearch, KNN_FINDNEAREST, layout)

    def FindNearestPos(self, Nsearch=64, layout=LAYOUT_BY_ID):
        """FindNearestPos(Int_t**,...) (KDFindNearest.cxx:448-450)."""
        return self._knn(Nsearch, KNN_FINDNEARESTPOS, layout)

    def FindNearestPhase(self, Nsearch=64, layout=LAYOUT_BY_ID):
        """FindNearestPhase(Int_t**,...) (KDFindNearest.cxx:454-456)."""
        return self._knn(Nsearch, KNN_FINDNEARESTPHASE, layout)

    def FindNearestNonPeriodic(self, Nsearch=64, layout=LAYOUT_BY_ID):
        """The search CalcDensity makes (KDCalcSmoothQuantities.cxx:227)."""
        return self._knn(Nsearch, KNN_NONPERIODIC, layout)

    def FindNearestVel(self, Nsearch=64, layout=LAYOUT_BY_ID):
        """FindNearestVel(Int_t**,...) on a TVEL tree (KDFindNearest.cxx:335-346, :451-453): the
        target search in velocity space, self and zero distances excluded."""
        if self.treetype != TVEL:
            raise NbxError(NBX_ERR_UNSUPPORTED, "FindNearestVel needs a TVEL tree")
        return self._knn(Nsearch, KNN_FINDNEAREST, layout)

    def FindNearestVelPoints(self, v, Nsearch=64):
        """FindNearestVel(Double_t *v, ...) (KDFindNearest.cxx:530-541) for many velocities v (m, 3), TVEL tree."""
        if self.treetype != TVEL:
            raise NbxError(NBX_ERR_UNSUPPORTED, "FindNearestVel needs a TVEL tree")
        return self.FindNearestPosPoints(v, Nsearch)

    def FindNearestPosPoints(self, x, Nsearch=64):
        """FindNearestPos(Double_t *x, ...) for many positions x (m, 3)."""
        x = np.asarray(x, dtype=np.float64)
        m = x.shape[0]
        q = np.ascontiguousarray(x.T)
        nn = np.zeros((m, Nsearch), dtype=np.int32)
        d2 = np.zeros((m, Nsearch), dtype=np.float64)
        _check(lib().nbx_knn_points(self._h, _p(q), m, Nsearch, _p(nn), _p(d2)))
        return nn, d2

    def FindNearestTargets(self, ids, Nsearch=64, variant=KNN_FINDNEAREST, layout=LAYOUT_BY_ID):
        """FindNearest(tt) (or the variant's per-particle call) for the particles with the given IDs;
        row r belongs to ids[r]."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        m = ids.shape[0]
        nn = np.zeros((m, Nsearch), dtype=np.int32)
        d2 = np.zeros((m, Nsearch), dtype=np.float64)
        _check(lib().nbx_knn_targets(self._h, _p(ids), m, Nsearch, variant, layout, _p(nn), _p(d2)))
        return nn, d2

    def FindNearestPhasePoints(self, x, v, Nsearch=64):
        """FindNearestPhase(Double_t *x, Double_t *v, ...) (KDFindNearest.cxx:528-538) for many phase
        points (x, v: (m, 3) each), TPHS tree; entries are IDs."""
        x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
        m = x.shape[0]
        if v.shape[0] != m:
            raise ValueError("x and v must have the same number of rows")
        qx, qv = np.ascontiguousarray(x.T), np.ascontiguousarray(v.T)
        nn = np.zeros((m, Nsearch), dtype=np.int32)
        d2 = np.zeros((m, Nsearch), dtype=np.float64)
        _check(lib().nbx_knn_phase_points(self._h, _p(qx), _p(qv), m, Nsearch, _p(nn), _p(d2)))
        return nn, d2

    # -------------------------------------------------------------- densities
    def CalcDensity(self, Nsmooth=64, return_h=False):
        """KDTree::CalcDensity (KDCalcSmoothQuantities.cxx:203-305); rho by ID."""
        rho = np.zeros(self.n, dtype=np.float64)
        h = np.zeros(self.n, dtype=np.float64) if return_h else None
        _check(lib().nbx_density(self._h, Nsmooth, DENS_SYMMETRIC, 0, _p(rho), _p(h)))
        return (rho, h) if return_h else rho

    def CalcDensityParticleAll(self, Nsmooth=64):
        """CalcDensityParticle(i, Nsmooth) for every i (KDCalcSmoothQuantities.cxx:768-844)."""
        rho = np.zeros(self.n, dtype=np.float64)
        h = np.zeros(self.n, dtype=np.float64)
        _check(lib().nbx_density(self._h, Nsmooth, DENS_GATHER, 0, _p(rho), _p(h)))
        return rho, h

    def CalcDensityPosition(self, x, Nsmooth=64, return_h=False):
        """CalcDensityPosition(x, Nsmooth) (KDCalcSmoothQuantities.cxx:1092-1149) at each row of x (m, 3);
        a single position (3,) gives a float."""
        x = np.asarray(x, dtype=np.float64)
        one = x.ndim == 1
        x = x.reshape(-1, 3)
        m = x.shape[0]
        q = np.ascontiguousarray(x.T)
        rho = np.zeros(m, dtype=np.float64)
        h = np.zeros(m, dtype=np.float64)
        _check(lib().nbx_density_points(self._h, _p(q), m, Nsmooth, _p(rho), _p(h)))
        if one:
            return (float(rho[0]), float(h[0])) if return_h else float(rho[0])
        return (rho, h) if return_h else rho

    def CalcVelDensityWithPhysDensity(self, Nsmooth=64, Nsearch=64, rho=None):
        """CalcVelDensityWithPhysDensity(Nsmooth, Nsearch, densityset) (KDCalcSmoothQuantities.cxx:393-476),
        keyed by ID.  rho (by ID) plays densityset = 1; None runs CalcDensity(Nsmooth) first, with
        Nsmooth > Nsearch taken as Nsearch before it, as the reference does (:405-409)."""
        Nsmooth = min(Nsmooth, Nsearch)
        if rho is None:
            rho = self.CalcDensity(Nsmooth)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.shape != (self.n,):
            raise ValueError("rho must have shape (n,)")
        out = np.zeros(self.n, dtype=np.float64)
        _check(lib().nbx_veldensity_phys(self._h, Nsmooth, Nsearch, _p(rho), _p(out)))
        return out

    def CalcVelDensityPosition(self, x, v, Nsmooth=64, Nsearch=64, return_h=False):
        """CalcVelDensityPosition(x, v, Nsmooth, Nsearch) (KDCalcSmoothQuantities.cxx:1150-1208) at each
        row of x, v (m, 3); a single point (3,) gives a float.  h is the velocity smoothing length."""
        x = np.asarray(x, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        one = x.ndim == 1
        x = x.reshape(-1, 3)
        v = v.reshape(-1, 3)
        if v.shape != x.shape:
            raise ValueError("x and v must have the same shape")
        m = x.shape[0]
        q, qv = np.ascontiguousarray(x.T), np.ascontiguousarray(v.T)
        rho = np.zeros(m, dtype=np.float64)
        h = np.zeros(m, dtype=np.float64)
        _check(lib().nbx_veldensity_points(self._h, _p(q), _p(qv), m, Nsmooth, Nsearch, _p(rho), _p(h)))
        if one:
            return (float(rho[0]), float(h[0])) if return_h else float(rho[0])
        return (rho, h) if return_h else rho

    def CalcSmoothVel(self, Nsmooth=64, rho=None):
        """CalcSmoothVel(Nsmooth, densityset) (KDCalcSmoothQuantities.cxx:480-540): smoothed velocities (n, 3)
        by ID.  rho: densities by ID to divide by (densityset = 1), or None to compute CalcDensity first."""
        return self._smooth(Nsmooth, rho, False)

    def CalcSmoothVelDisp(self, Nsmooth=64, rho=None, mean=None):
        """CalcSmoothVelDisp (:542-614): (mean (n, 3), dispersion (n, 3, 3)) by ID.  mean: the means to
        take (meanvelset = 1), or None to compute CalcSmoothVel first."""
        return self._smooth(Nsmooth, rho, True, mean)

    def CalcSmoothVelMoments(self, Nsmooth=64, rho=None, mean=None, disp=None):
        """CalcSmoothVelSkew (:617-690) and CalcSmoothVelKurtosis (:692-767): (skew, kurt), each (n, 3) by ID.
        mean (n, 3) / disp (n, 3, 3): the means and dispersions to take (meanvelset / veldispset = 1), or None
        to compute them."""
        skew = np.zeros((self.n, 3), dtype=np.float64)
        kurt = np.zeros((self.n, 3), dtype=np.float64)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        mi = None if mean is None else np.ascontiguousarray(mean, dtype=np.float64).reshape(self.n, 3)
        di = None if disp is None else np.ascontiguousarray(disp, dtype=np.float64).reshape(self.n, 9)
        _check(lib().nbx_smooth_vel_moments(self._h, Nsmooth, _p(r), _p(mi), _p(di), _p(skew), _p(kurt)))
        return skew, kurt

    def CalcSmoothVelSkew(self, Nsmooth=64, rho=None, mean=None, disp=None):
        """CalcSmoothVelSkew (KDCalcSmoothQuantities.cxx:617-690): (n, 3) by ID."""
        return self.CalcSmoothVelMoments(Nsmooth, rho, mean, disp)[0]

    def CalcSmoothVelKurtosis(self, Nsmooth=64, rho=None, mean=None, disp=None):
        """CalcSmoothVelKurtosis (KDCalcSmoothQuantities.cxx:692-767): (n, 3) by ID."""
        return self.CalcSmoothVelMoments(Nsmooth, rho, mean, disp)[1]

    def CalcSmoothLocalMeanAll(self, weight, Nsmooth=64, rho=None):
        """CalcSmoothLocalMeanParticle(i, weight, Nsmooth) (KDCalcSmoothQuantities.cxx:1539-1577) for every
        particle: (n,) by ID.  weight: one value per particle (by ID); rho: the densities to divide by, or
        None to compute CalcDensity first."""
        return self._local(weight, Nsmooth, rho, None)[0]

    def CalcSmoothLocalDispAll(self, weight, mean, Nsmooth=64, rho=None):
        """CalcSmoothLocalDispParticle(i, weight, mean[i], Nsmooth) (:1579-1617) for every particle."""
        return self._local(weight, Nsmooth, rho, mean)[1]

    def CalcSmoothLocalMeanPosition(self, x, weight, Nsmooth=64, rho=None):
        """CalcSmoothLocalMeanPosition(x, weight, Nsmooth) (:1619-1655) at every row of x (m, 3)."""
        return self._local(weight, Nsmooth, rho, None, x)[0]

    def CalcSmoothLocalDispPosition(self, x, weight, mean, Nsmooth=64, rho=None):
        """CalcSmoothLocalDispPosition(x, weight, mean, Nsmooth) (:1657-1693) at every row of x (m, 3); mean:
        one value per row."""
        return self._local(weight, Nsmooth, rho, mean, x)[1]

    def _local(self, weight, k, rho, mean_in, x=None):
        w = np.ascontiguousarray(weight, dtype=np.float64)
        if w.shape != (self.n,):
            raise ValueError("weight must have one value per particle")
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        mi = None if mean_in is None else np.ascontiguousarray(mean_in, dtype=np.float64)
        if x is None:
            mean = np.zeros(self.n)
            disp = np.zeros(self.n) if mi is not None else None
            _check(lib().nbx_smooth_local(self._h, k, _p(w), _p(r), _p(mi), _p(mean), _p(disp)))
        else:
            x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
            mm = x.shape[0]
            q = np.ascontiguousarray(x.T)
            mean = np.zeros(mm)
            disp = np.zeros(mm) if mi is not None else None
            _check(lib().nbx_smooth_local_points(self._h, _p(q), mm, k, _p(w), _p(r), _p(mi), _p(mean), _p(disp)))
        return mean, disp

    def _smooth(self, k, rho, with_disp, mean_in=None):
        mean = np.zeros((self.n, 3), dtype=np.float64)
        disp = np.zeros((self.n, 3, 3), dtype=np.float64) if with_disp else None
        r = mi = None
        if rho is not None:
            r = np.ascontiguousarray(rho, dtype=np.float64)
            if r.shape != (self.n,):
                raise ValueError("rho must have one value per particle")
        if mean_in is not None:
            mi = np.ascontiguousarray(mean_in, dtype=np.float64)
            if mi.shape != (self.n, 3):
                raise ValueError("mean must be (n, 3)")
            mean[:] = mi
        _check(lib().nbx_smooth_vel(self._h, k, _p(r), _p(mi), None if mi is not None else _p(mean), _p(disp)))
        return (mean, disp) if with_disp else mean

    def CalcVelDensity(self, Nsmooth=64, Nsearch=64):
        """KDTree::CalcVelDensity (KDCalcSmoothQuantities.cxx:309-389); result by ID."""
        rho = np.zeros(self.n, dtype=np.float64)
        _check(lib().nbx_density(self._h, Nsmooth, DENS_VEL, Nsearch, _p(rho), None))
        return rho

    # -------------------------------------------------------------- ball / FOF
    def SearchBallPosTagged(self, x, fdist2, cap=None):
        """SearchBallPosTagged(Double_t *x, fdist2, tagged) (KDFindNearest.cxx:628-636); IDs, unordered."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))
        cap = self.n * 8 if cap is None else cap
        out = np.zeros(max(cap, 1), dtype=np.int32)
        nout = ctypes.c_int64()
        _check(lib().nbx_ball_points(self._h, _p(x), float(fdist2), _p(out), cap, ctypes.byref(nout)))
        return out[: min(nout.value, cap)].copy()

    def SearchBallPosTaggedMany(self, x, fdist2):
        """SearchBallPosTagged(x, fdist2) for every row of x (m, 3) in one call: (offsets (m+1,), ids, d2),
        row i's IDs being ids[offsets[i]:offsets[i+1]] (unordered)."""
        x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
        m = x.shape[0]
        q = np.ascontiguousarray(x.T)
        off = np.zeros(m + 1, dtype=np.int64)
        _check(lib().nbx_ball_points_many(self._h, _p(q), m, float(fdist2), _p(off), None, None, 0))
        total = int(off[-1])
        ids = np.zeros(max(total, 1), dtype=np.int32)
        d2 = np.zeros(max(total, 1), dtype=np.float64)
        if total:
            _check(lib().nbx_ball_points_many(self._h, _p(q), m, float(fdist2), _p(off), _p(ids), _p(d2), total))
        return off, ids[:total], d2[:total]

    def SearchBallPos(self, x, fdist2, imark, nn, dist2, exclude_id=-1):
        """SearchBallPos(x, fdist2, imark, nn, dist2) (KDFindNearest.cxx:575-582): nn[id] = imark and
        dist2[id] = d2 for every particle with d2 < fdist2 (ID-indexed arrays of n, updated in place);
        exclude_id skips one particle (the target of the SearchBallPos(tt, ...) form)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))
        if nn.dtype != np.int32 or dist2.dtype != np.float64 or nn.shape != (self.n,) or dist2.shape != (self.n,):
            raise ValueError("nn must be int32[n] and dist2 float64[n]")
        if not (nn.flags.c_contiguous and dist2.flags.c_contiguous):
            raise ValueError("nn and dist2 must be contiguous")
        _check(lib().nbx_ball_mark(self._h, _p(x), int(exclude_id), float(fdist2), int(imark), _p(nn), _p(dist2)))

    def FOF(self, fdist, minnum=8):
        """KDTree::FOF (KDFOF.cxx:29-153): canonical labels by ID and the group count."""
        lab = np.zeros(self.n, dtype=np.int32)
        ng = ctypes.c_int64()
        _check(lib().nbx_fof(self._h, float(fdist), minnum, _p(lab), ctypes.byref(ng)))
        return lab, ng.value

    def FOFCriterion(self, comparator, params, minnum=8):
        """KDTree::FOFCriterion(cmp, params, numgroups, minnum) (KDFOF.cxx:157-265) with cmp = FOF3d
        (FOF_CMP_3D) or FOF6d (FOF_CMP_6D), FOFFunc.h:30-55.  params as the reference's (up to 8 values:
        [0] treetype, [1]/[2] search scales, [6]/[7] comparator scales).  Canonical labels by ID and the
        group count."""
        p = np.zeros(8, dtype=np.float64)
        v = np.asarray(params, dtype=np.float64).ravel()[:8]
        p[: v.size] = v
        lab = np.zeros(self.n, dtype=np.int32)
        ng = ctypes.c_int64()
        _check(lib().nbx_fof_criterion(self._h, int(comparator), _p(p), minnum, _p(lab), ctypes.byref(ng)))
        return lab, ng.value


def fof_lists(labels_by_id, slot_to_id, ngroups):
    """FOF's pHead/pNext/pTail/pLen (KDFOF.cxx:29-31) from canonical labels, by tree index (nbx_fof_lists):
    members of each group linked in slot order; host only."""
    lab = np.ascontiguousarray(labels_by_id, dtype=np.int32)
    order = np.ascontiguousarray(slot_to_id, dtype=np.int32)
    n = lab.shape[0]
    head, nxt, tail, ln = (np.zeros(n, dtype=np.int32) for _ in range(4))
    _check(lib().nbx_fof_lists(_p(lab), _p(order), n, int(ngroups), _p(head), _p(nxt), _p(tail), _p(ln)))
    return head, nxt, tail, ln


def prof_enable(on=True):
    lib().nbx_prof_enable(1 if on else 0)


def prof_reset():
    lib().nbx_prof_reset()


def prof_get(name):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    lib().nbx_prof_get(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value

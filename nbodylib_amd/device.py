"""Device-resident entry points over the C ABI (torch tensors as HBM buffers).

torch supplies device memory and the stream; all compute runs in libnbx.so.
Positions/velocities are (3, n) float64 CUDA tensors (SoA), masses (n,).

SPHDensity is the differentiation form: a torch.autograd.Function whose
forward is the symmetric SPH density over a frozen neighbour list and whose
backward is the hand-written adjoint kernel (nbx_sph_density_adjoint_dev).
"""
import ctypes

import torch

from . import NbxError, _check, lib, params, FORM_PLAIN, KEPAN, TPHYS, LAYOUT_BY_SLOT, KNN_FINDNEAREST


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(t, name, dtype, shape=None):
    if t is None:
        return None
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous CUDA {dtype} tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
    return t


class DeviceTree:
    """KDTree built from HBM-resident tensors (nbx_tree_create_dev)."""

    def __init__(self, pos, mass=None, vel=None, bucket_size=16, treetype=TPHYS, kerntype=KEPAN, kernres=1000,
                 aniso=0, period=None, form=FORM_PLAIN, recheck_rel=1e-5):
        n = pos.shape[1]
        self.n = int(n)
        _need(pos, "pos", torch.float64, (3, n))
        _need(vel, "vel", torch.float64, (3, n))
        _need(mass, "mass", torch.float64, (n,))
        self._keep = (pos, mass, vel)
        self._params = params(bucket_size, treetype, kerntype, kernres, 0, aniso, 0, period, form, recheck_rel)
        h = ctypes.c_void_p()
        _check(lib().nbx_tree_create_dev(ctypes.byref(self._params), _ptr(pos), _ptr(vel), _ptr(mass), self.n,
                                         _stream(), ctypes.byref(h)))
        self._h = h
        self._destroy = lib().nbx_tree_destroy
        self._keep = None  # the tree holds its own sorted copies

    def __del__(self):
        # the destroy function is bound at creation: at interpreter exit the
        # module's globals may already be gone
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self._h = None

    def order(self):
        """slot -> ID permutation (int32 CUDA tensor): tree slot s holds input particle order()[s]."""
        import numpy as np

        o = np.empty(self.n, dtype=np.int32)
        _check(lib().nbx_tree_order(self._h, ctypes.c_void_p(o.ctypes.data)))
        return torch.from_numpy(o).to(torch.device("cuda", torch.cuda.current_device()))

    def last_fallbacks(self):
        v = ctypes.c_int64()
        _check(lib().nbx_tree_last_fallbacks(self._h, ctypes.byref(v)))
        return v.value

    def knn(self, k, nn_out, d2_out=None, variant=KNN_FINDNEAREST, layout=LAYOUT_BY_SLOT):
        _need(nn_out, "nn_out", torch.int32, (self.n, k))
        _need(d2_out, "d2_out", torch.float64, (self.n, k))
        _check(lib().nbx_knn_dev(self._h, k, variant, layout, _ptr(nn_out), _ptr(d2_out), _stream()))

    def knn_density(self, k, nn_out, rho_out, h_out=None):
        """Fused kNN + CalcDensity (neighbour slots by slot, rho/h by ID)."""
        _need(nn_out, "nn_out", torch.int32, (self.n, k))
        _need(rho_out, "rho_out", torch.float64, (self.n,))
        _need(h_out, "h_out", torch.float64, (self.n,))
        _check(lib().nbx_knn_density_dev(self._h, k, _ptr(nn_out), _ptr(rho_out), _ptr(h_out), _stream()))

    def sph_density(self, k, rho_out, h_out=None, block_rows=1 << 22):
        """Differentiation-form forward over the tree's own k-nearest lists, recomputed per block of
        block_rows tree slots (nbx_sph_density_fwd_tree_dev); rho / h by ID."""
        _need(rho_out, "rho_out", torch.float64, (self.n,))
        _need(h_out, "h_out", torch.float64, (self.n,))
        _check(lib().nbx_sph_density_fwd_tree_dev(self._h, k, int(block_rows), _ptr(rho_out), _ptr(h_out),
                                                  _stream()))

    def sph_density_adjoint(self, k, g, dpos_out, dmass_out, block_rows=1 << 22):
        """Its adjoint (nbx_sph_density_adjoint_tree_dev): g = dL/drho by ID; dpos (3, n), dmass (n,) by ID."""
        _need(g, "g", torch.float64, (self.n,))
        _need(dpos_out, "dpos_out", torch.float64, (3, self.n))
        _need(dmass_out, "dmass_out", torch.float64, (self.n,))
        _check(lib().nbx_sph_density_adjoint_tree_dev(self._h, k, int(block_rows), _ptr(g), _ptr(dpos_out),
                                                      _ptr(dmass_out), _stream()))

    def density(self, k, rho_out, h_out=None, variant=0, nsearch=0):
        _need(rho_out, "rho_out", torch.float64, (self.n,))
        _need(h_out, "h_out", torch.float64, (self.n,))
        _check(lib().nbx_density_dev(self._h, k, variant, nsearch, _ptr(rho_out), _ptr(h_out), _stream()))

    def fof(self, fdist, minnum, labels_out):
        _need(labels_out, "labels_out", torch.int32, (self.n,))
        ng = ctypes.c_int64()
        _check(lib().nbx_fof_dev(self._h, float(fdist), minnum, _ptr(labels_out), ctypes.byref(ng), _stream()))
        return ng.value

    def fof_criterion(self, comparator, params, minnum, labels_out):
        """FOFCriterion with FOF3d / FOF6d (nbx_fof_criterion_dev); params as the reference's (8 values)."""
        import numpy as np

        _need(labels_out, "labels_out", torch.int32, (self.n,))
        p = np.zeros(8, dtype=np.float64)
        v = np.asarray(params, dtype=np.float64).ravel()[:8]
        p[: v.size] = v
        ng = ctypes.c_int64()
        _check(lib().nbx_fof_criterion_dev(self._h, int(comparator), p.ctypes.data_as(ctypes.c_void_p), minnum,
                                           _ptr(labels_out), ctypes.byref(ng), _stream()))
        return ng.value


def _check_list(nbr, n):
    if nbr.dtype != torch.int32 or not nbr.is_cuda or nbr.dim() != 2 or nbr.shape[0] != n:
        raise ValueError("nbr must be an (n, k) int32 CUDA tensor")
    lo, hi = int(nbr.min()), int(nbr.max())
    if lo < 0 or hi >= n:
        raise ValueError("neighbour indices out of range")


def sph_density(pos, mass, nbr):
    """Symmetric SPH density over a frozen list (analytic cubic spline)."""
    n = pos.shape[1]
    _need(pos, "pos", torch.float64, (3, n))
    _need(mass, "mass", torch.float64, (n,))
    rho = torch.empty(n, dtype=torch.float64, device=pos.device)
    h = torch.empty(n, dtype=torch.float64, device=pos.device)
    _check(lib().nbx_sph_density_fwd_dev(_ptr(pos), _ptr(mass), _ptr(nbr), n, nbr.shape[1], _ptr(rho), _ptr(h),
                                         _stream()))
    return rho, h


class SPHDensity(torch.autograd.Function):
    """rho = SPHDensity.apply(pos (3,n) f64, mass (n,) f64, nbr (n,k) int32).

    Backward runs the hand-written adjoint (d rho / d pos, d rho / d mass)
    over the same frozen list; the list itself is not differentiated.
    """

    @staticmethod
    def forward(ctx, pos, mass, nbr):
        pos = pos.contiguous()
        mass = mass.contiguous()
        nbr = nbr.contiguous()
        _check_list(nbr, pos.shape[1])
        rho, _ = sph_density(pos, mass, nbr)
        ctx.save_for_backward(pos, mass, nbr)
        return rho

    @staticmethod
    def backward(ctx, g):
        pos, mass, nbr = ctx.saved_tensors
        n, k = nbr.shape
        g = g.contiguous().to(torch.float64)
        dpos = torch.empty_like(pos)
        dmass = torch.empty_like(mass)
        _check(lib().nbx_sph_density_adjoint_dev(_ptr(pos), _ptr(mass), _ptr(nbr), n, k, _ptr(g), _ptr(dpos),
                                                 _ptr(dmass), _stream()))
        return dpos, dmass, None


__all__ = ["DeviceTree", "SPHDensity", "sph_density", "NbxError"]

"""Host-side mirror of the reference's triangle siblings over the HIP C ABI.

Names, argument order and meaning follow the reference (``namespace sim``):
``linear_tri2dmesh_{neohookean,stvk,corotational,arap}_{q,dq,dq2}(V, E, q, dphidX, volume, params[, callback])``,
``linear_tri2dmesh_dphi_dX(V, E)``, ``linear_tri2dmesh_mass_matrix(V, E, density, volume)``,
``linear_tri3dmesh_dphi_dX(V, E)`` and ``linear_tri3dmesh_stvk_dq2(V, E, q, dphidX, N, n, dndq, volume, params[,
callback])``.  Arrays use the reference's Eigen (column-major) layouts (dphidX nT x 6 or nT x 9 with
Phi(m, j) at column m + 3j; q with 2 or 3 values per vertex, interleaved); the sparse matrices come
back as ``scipy.sparse.csc_matrix`` with full dim x dim blocks per vertex pair, explicit zeros kept
(the reference's pattern).  The element callback is ``tetmesh.simple_psd_fix(tol)``.
"""
import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check
from .tetmesh import ARAP, COROTATIONAL, NEOHOOKEAN, STVK, _colmajor, _psd_struct, _ptr, simple_psd_fix  # noqa: F401


def linear_tri2dmesh_dphi_dX(V, E):
    """nT x 6.  Reference: src/linear_tri2dmesh_dphi_dX.cpp:5-16, src/linear_tri2d_dphi_dX.cpp:5-17."""
    lib = _lib.load()
    V, E = _colmajor(V, np.float64), _colmajor(E, np.int32)
    dX = np.zeros((E.shape[0], 6), order="F")
    check(lib.bh_linear_tri2dmesh_dphi_dX(V.shape[0], _ptr(V), E.shape[0], _ptr(E), _ptr(dX)),
          "bh_linear_tri2dmesh_dphi_dX")
    return dX


def linear_tri3dmesh_dphi_dX(V, E):
    """nT x 9.  Reference: src/linear_tri3dmesh_dphi_dX.cpp:5-19, src/linear_tri3d_dphi_dX.cpp:5-19."""
    lib = _lib.load()
    V, E = _colmajor(V, np.float64), _colmajor(E, np.int32)
    dX = np.zeros((E.shape[0], 9), order="F")
    check(lib.bh_linear_tri3dmesh_dphi_dX(V.shape[0], _ptr(V), E.shape[0], _ptr(E), _ptr(dX)),
          "bh_linear_tri3dmesh_dphi_dX")
    return dX


class TriMesh:
    """Pattern and incidence lists of a triangle mesh with dim (2 or 3) DOFs per vertex, built once."""

    def __init__(self, dim, nV, E):
        import ctypes

        self.lib = _lib.load()
        self.dim, self.nV = int(dim), int(nV)
        E = _colmajor(E, np.int32)
        self.nT = E.shape[0]
        h = ctypes.c_void_p()
        check(self.lib.bh_trimesh_create(self.dim, self.nV, self.nT, _ptr(E), ctypes.byref(h)), "bh_trimesh_create")
        self.handle = h
        n = ctypes.c_int64()
        check(self.lib.bh_trimesh_nnz(h, ctypes.byref(n)), "bh_trimesh_nnz")
        self.nnz = int(n.value)
        self._pattern = None

    def close(self):
        if getattr(self, "handle", None):
            self.lib.bh_trimesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        if self._pattern is None:
            rowptr = np.zeros(self.dim * self.nV + 1, np.int32)
            colind = np.zeros(self.nnz, np.int32)
            check(self.lib.bh_trimesh_pattern(self.handle, _ptr(rowptr), _ptr(colind)), "bh_trimesh_pattern")
            self._pattern = (rowptr, colind)
        return self._pattern

    def _csc(self, vals):
        rowptr, colind = self.pattern()
        n = self.dim * self.nV
        # the pattern is symmetric and so is every element matrix: CSR arrays read as CSC give H
        return sp.csc_matrix((vals, colind, rowptr), shape=(n, n))

    @staticmethod
    def _args(q, dphidX, volume, params):
        return (np.ascontiguousarray(q, dtype=np.float64), _colmajor(dphidX, np.float64),
                np.ascontiguousarray(volume, dtype=np.float64).ravel(), _colmajor(params, np.float64))

    def q(self, model, q, dphidX, volume, params):
        import ctypes

        q, dX, vol, p = self._args(q, dphidX, volume, params)
        e = ctypes.c_double()
        check(self.lib.bh_linear_tri2dmesh_q(self.handle, model, _ptr(q), _ptr(dX), _ptr(vol), _ptr(p), ctypes.byref(e)),
              "bh_linear_tri2dmesh_q")
        return e.value

    def dq(self, model, q, dphidX, volume, params):
        q, dX, vol, p = self._args(q, dphidX, volume, params)
        g = np.zeros(self.dim * self.nV)
        check(self.lib.bh_linear_tri2dmesh_dq(self.handle, model, _ptr(q), _ptr(dX), _ptr(vol), _ptr(p), _ptr(g)),
              "bh_linear_tri2dmesh_dq")
        return g

    def dq2_values(self, model, q, dphidX, volume, params, callback=None):
        import ctypes

        q, dX, vol, p = self._args(q, dphidX, volume, params)
        vals = np.zeros(self.nnz)
        psd = _psd_struct(callback)
        check(self.lib.bh_linear_tri2dmesh_dq2(self.handle, model, _ptr(q), _ptr(dX), _ptr(vol), _ptr(p),
                                               ctypes.byref(psd) if psd is not None else None, _ptr(vals)),
              "bh_linear_tri2dmesh_dq2")
        return vals

    def dq2(self, model, q, dphidX, volume, params, callback=None):
        return self._csc(self.dq2_values(model, q, dphidX, volume, params, callback))

    def mass_matrix(self, density, volume):
        vals = np.zeros(self.nnz)
        rho = np.ascontiguousarray(density, dtype=np.float64).ravel()
        vol = np.ascontiguousarray(volume, dtype=np.float64).ravel()
        check(self.lib.bh_linear_tri2dmesh_mass_matrix(self.handle, _ptr(rho), _ptr(vol), _ptr(vals)),
              "bh_linear_tri2dmesh_mass_matrix")
        return self._csc(vals)

    def stvk_shell_dq2(self, q, dphidX, N, n, dndq, volume, params, callback=None):
        import ctypes

        q, dX, vol, p = self._args(q, dphidX, volume, params)
        N, n, dndq = _colmajor(N, np.float64), _colmajor(n, np.float64), _colmajor(dndq, np.float64)
        vals = np.zeros(self.nnz)
        psd = _psd_struct(callback)
        check(self.lib.bh_linear_tri3dmesh_stvk_dq2(self.handle, _ptr(q), _ptr(dX), _ptr(N), _ptr(n), _ptr(dndq),
                                                    _ptr(vol), _ptr(p), ctypes.byref(psd) if psd is not None else None,
                                                    _ptr(vals)),
              "bh_linear_tri3dmesh_stvk_dq2")
        return self._csc(vals)


def _dp(t):
    """device pointer of a torch tensor (or None)"""
    import ctypes

    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import ctypes

    return None if stream is None else ctypes.c_void_p(stream.cuda_stream)


class TriMeshDevice(TriMesh):
    """The same calls on device arrays (torch tensors, FP64, the reference's column-major layouts
    flattened: dphidX.ravel(order="F") etc.); each call synchronises its stream."""

    def q_device(self, model, q, dphidX, volume, params, out, stream=None):
        check(self.lib.bh_linear_tri2dmesh_q_device(self.handle, model, _dp(q), _dp(dphidX), _dp(volume), _dp(params),
                                                    _dp(out), _stream(stream)), "bh_linear_tri2dmesh_q_device")

    def dq_device(self, model, q, dphidX, volume, params, out, stream=None):
        check(self.lib.bh_linear_tri2dmesh_dq_device(self.handle, model, _dp(q), _dp(dphidX), _dp(volume),
                                                     _dp(params), _dp(out), _stream(stream)),
              "bh_linear_tri2dmesh_dq_device")

    def dq2_device(self, model, q, dphidX, volume, params, out, callback=None, stream=None):
        import ctypes

        psd = _psd_struct(callback)
        check(self.lib.bh_linear_tri2dmesh_dq2_device(self.handle, model, _dp(q), _dp(dphidX), _dp(volume),
                                                      _dp(params), ctypes.byref(psd) if psd is not None else None,
                                                      _dp(out), _stream(stream)), "bh_linear_tri2dmesh_dq2_device")

    def stvk_shell_dq2_device(self, q, dphidX, N, n, dndq, volume, params, out, callback=None, stream=None):
        import ctypes

        psd = _psd_struct(callback)
        check(self.lib.bh_linear_tri3dmesh_stvk_dq2_device(self.handle, _dp(q), _dp(dphidX), _dp(N), _dp(n), _dp(dndq),
                                                           _dp(volume), _dp(params),
                                                           ctypes.byref(psd) if psd is not None else None, _dp(out),
                                                           _stream(stream)), "bh_linear_tri3dmesh_stvk_dq2_device")


def _make(model):
    def q_fn(V, E, q, dphidX, volume, params):
        return TriMesh(2, np.asarray(V).shape[0], E).q(model, q, dphidX, volume, params)

    def dq_fn(V, E, q, dphidX, volume, params):
        return TriMesh(2, np.asarray(V).shape[0], E).dq(model, q, dphidX, volume, params)

    def dq2_fn(V, E, q, dphidX, volume, params, callback=None):
        return TriMesh(2, np.asarray(V).shape[0], E).dq2(model, q, dphidX, volume, params, callback)

    return q_fn, dq_fn, dq2_fn


# src/linear_tri2dmesh_{neohookean,stvk,corotational,arap}_{q,dq,dq2}.cpp
linear_tri2dmesh_neohookean_q, linear_tri2dmesh_neohookean_dq, linear_tri2dmesh_neohookean_dq2 = _make(NEOHOOKEAN)
linear_tri2dmesh_stvk_q, linear_tri2dmesh_stvk_dq, linear_tri2dmesh_stvk_dq2 = _make(STVK)
linear_tri2dmesh_corotational_q, linear_tri2dmesh_corotational_dq, linear_tri2dmesh_corotational_dq2 = _make(COROTATIONAL)
linear_tri2dmesh_arap_q, linear_tri2dmesh_arap_dq, linear_tri2dmesh_arap_dq2 = _make(ARAP)


def linear_tri2dmesh_mass_matrix(V, E, density, volume):
    """Reference: src/linear_tri2dmesh_mass_matrix.cpp, src/linear_tri2d_mass_matrix.cpp:5-29."""
    return TriMesh(2, np.asarray(V).shape[0], E).mass_matrix(density, volume)


def linear_tri3dmesh_stvk_dq2(V, E, q, dphidX, N, n, dndq, volume, params, callback=None):
    """Reference: src/linear_tri3dmesh_stvk_dq2.cpp:5-30, src/linear_tri3d_stvk_dq2.cpp:6-40."""
    return TriMesh(3, np.asarray(V).shape[0], E).stvk_shell_dq2(q, dphidX, N, n, dndq, volume, params, callback)

"""Host-side mirror of the reference's linear-tetmesh interface over the HIP C ABI.

Names, argument order and meaning follow the reference (``namespace sim``):
``linear_tetmesh_{neohookean,stvk,corotational,arap}_{q,dq,dq2}(V, E, q, dphidX, volume, params[, callback])``
and ``linear_tetmesh_generic_elastic_*``.  Arrays use the reference's Eigen (column-major)
layouts; the sparse Hessian comes back as a ``scipy.sparse.csc_matrix`` whose arrays are exactly
``Eigen::SparseMatrix<double>``'s (outer = column pointers, inner = sorted row indices).

The only element-matrix callback the GPU path supports is the one every in-repo caller uses,
``simple_psd_fix(tol)`` (src/simple_psd_fix.cpp:5-21); pass ``simple_psd_fix(1e-3)`` where the
reference takes ``[](auto &a){ sim::simple_psd_fix(a, 1e-3); }``.  Other callables raise.
"""
import ctypes
import hashlib

import numpy as np

from . import _lib
from ._lib import check

NEOHOOKEAN, STVK, COROTATIONAL, ARAP, NEOHOOKEAN_LEGACY, NEOHOOKEAN_MUSCLE = 0, 1, 2, 3, 4, 5
MODEL_PER_TET = 255


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _colmajor(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class simple_psd_fix:
    """marker for the reference's simple_psd_fix(A, tol) element callback (default tol 1e-6,
    include/simple_psd_fix.h:11; the MATLAB wrappers pass 1e-3)."""

    def __init__(self, tol=1e-6):
        self.tol = float(tol)


def _psd_struct(callback):
    if callback is None:
        return None
    if isinstance(callback, simple_psd_fix):
        return _lib.BhPsd(1, callback.tol)
    raise TypeError("the device entry points take simple_psd_fix(tol); other element callbacks go through "
                    "TetMesh.dq2 / the linear_tetmesh_*_dq2 functions (host callback path)")


def _element_callback(callback):
    """a Python callable f(H) that changes the 12x12 element matrix H (a numpy view, column-major as the
    reference's Eigen::Matrix12d) in place, wrapped as a bh_element_callback"""
    def trampoline(ptr, element, user):
        H = np.ctypeslib.as_array(ptr, shape=(144,)).reshape(12, 12, order="F")
        callback(H)

    return _lib.ELEMENT_CALLBACK(trampoline)


def linear_tetmesh_dphi_dX(V, E):
    """(dphidX nT x 12, volume nT).  Reference: src/linear_tetmesh_dphi_dX.cpp:5-18 and igl::volume."""
    lib = _lib.load()
    V = _colmajor(V, np.float64)
    E = _colmajor(E, np.int32)
    nV, nT = V.shape[0], E.shape[0]
    dX = np.zeros((nT, 12), order="F")
    vol = np.zeros(nT)
    st = lib.bh_linear_tetmesh_dphi_dX(nV, _ptr(V), nT, _ptr(E), _ptr(dX), _ptr(vol))
    if st not in (_lib.BH_OK, _lib.BH_NONFINITE):
        check(st, "bh_linear_tetmesh_dphi_dX")
    return dX, vol


class TetMesh:
    """Per-mesh device state (pattern, incidence lists, element geometry), built once."""

    def __init__(self, nV, E, dphidX, volume, V=None):
        """V (rest positions, nV x 3), optional: lets the FP32 precision form evaluate stresses from
        displacement differences (bh_mesh_set_rest_positions)."""
        lib = _lib.load()
        E = _colmajor(E, np.int32)
        dphidX = _colmajor(dphidX, np.float64)
        volume = np.ascontiguousarray(volume, dtype=np.float64)
        self.nV, self.nT = int(nV), int(E.shape[0])
        h = ctypes.c_void_p()
        check(lib.bh_mesh_create(self.nV, self.nT, _ptr(E), _ptr(dphidX), _ptr(volume), ctypes.byref(h)),
              "bh_mesh_create")
        self._h = h
        self._lib = lib
        nnz = ctypes.c_int64()
        check(lib.bh_mesh_nnz(h, ctypes.byref(nnz)), "bh_mesh_nnz")
        self.nnz = int(nnz.value)
        self._pattern = None
        if V is not None:
            Vc = _colmajor(np.asarray(V, dtype=np.float64).reshape(self.nV, 3), np.float64)
            check(lib.bh_mesh_set_rest_positions(h, _ptr(Vc)), "bh_mesh_set_rest_positions")

    @classmethod
    def from_device(cls, nV, E, V=None, dphidX=None, volume=None):
        """the same mesh built on the device (bh_mesh_create_device) from torch tensors on cuda:0: E
        (nT x 4 column-major int32, flattened), and either dphidX (nT x 12 column-major) and volume, or
        V (nV x 3 column-major) to compute them on the device"""
        self = cls.__new__(cls)
        lib = _lib.load()
        dp = cls._dp
        nT = E.numel() // 4
        h = ctypes.c_void_p()
        check(lib.bh_mesh_create_device(int(nV), int(nT), dp(E), dp(V), dp(dphidX), dp(volume), ctypes.byref(h)),
              "bh_mesh_create_device")
        self._h, self._lib, self.nV, self.nT = h, lib, int(nV), int(nT)
        nnz = ctypes.c_int64()
        check(lib.bh_mesh_nnz(h, ctypes.byref(nnz)), "bh_mesh_nnz")
        self.nnz = int(nnz.value)
        self._pattern = None
        return self

    @property
    def handle(self):
        return self._h

    def trim(self):
        """free the work buffers the mesh keeps between calls (remade on next use): bh_mesh_trim"""
        if self._h:
            _lib.check(self._lib.bh_mesh_trim(self._h), "bh_mesh_trim")

    def close(self):
        if self._h:
            self._lib.bh_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stretch_reference_layout(self, on):
        o = _lib.BhElementOpts(1 if on else 0)
        check(self._lib.bh_mesh_set_element_opts(self._h, ctypes.byref(o)), "bh_mesh_set_element_opts")

    def pattern(self):
        """(rowptr int32[3nV+1], colind int32[nnz]) of the reference pattern (CSR == CSC)."""
        if self._pattern is None:
            rowptr = np.zeros(3 * self.nV + 1, dtype=np.int32)
            colind = np.zeros(self.nnz, dtype=np.int32)
            check(self._lib.bh_mesh_pattern(self._h, _ptr(rowptr), _ptr(colind)), "bh_mesh_pattern")
            self._pattern = (rowptr, colind)
        return self._pattern

    # ---- host arrays -------------------------------------------------------------------------------
    def q(self, model, q, params, model_ids=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        params = _colmajor(params, np.float64)
        out = ctypes.c_double()
        if model_ids is not None:
            ids = np.ascontiguousarray(model_ids, dtype=np.uint8)
            check(self._lib.bh_linear_tetmesh_generic_elastic_q(self._h, _ptr(ids), _ptr(q), _ptr(params),
                                                                ctypes.byref(out)), "generic_elastic_q")
        else:
            check(self._lib.bh_linear_tetmesh_q(self._h, model, _ptr(q), _ptr(params), ctypes.byref(out)),
                  "bh_linear_tetmesh_q")
        return out.value

    def dq(self, model, q, params, model_ids=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        params = _colmajor(params, np.float64)
        g = np.zeros(3 * self.nV)
        if model_ids is not None:
            ids = np.ascontiguousarray(model_ids, dtype=np.uint8)
            check(self._lib.bh_linear_tetmesh_generic_elastic_dq(self._h, _ptr(ids), _ptr(q), _ptr(params), _ptr(g)),
                  "generic_elastic_dq")
        else:
            check(self._lib.bh_linear_tetmesh_dq(self._h, model, _ptr(q), _ptr(params), _ptr(g)),
                  "bh_linear_tetmesh_dq")
        return g

    def dq2_values(self, model, q, params, callback=None, model_ids=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        params = _colmajor(params, np.float64)
        vals = np.zeros(self.nnz)
        if callback is not None and not isinstance(callback, simple_psd_fix):
            # an arbitrary element callback: element matrices made on the device, changed on the host
            cb = _element_callback(callback)
            ids = np.ascontiguousarray(model_ids, dtype=np.uint8) if model_ids is not None else None
            check(self._lib.bh_linear_tetmesh_dq2_callback(self._h, MODEL_PER_TET if ids is not None else model,
                                                           _ptr(ids) if ids is not None else None, _ptr(q),
                                                           _ptr(params), cb, None, _ptr(vals)),
                  "bh_linear_tetmesh_dq2_callback")
            return vals
        psd = _psd_struct(callback)
        pp = ctypes.byref(psd) if psd is not None else None
        if model_ids is not None:
            ids = np.ascontiguousarray(model_ids, dtype=np.uint8)
            check(self._lib.bh_linear_tetmesh_generic_elastic_dq2(self._h, _ptr(ids), _ptr(q), _ptr(params), pp,
                                                                  _ptr(vals)), "generic_elastic_dq2")
        else:
            check(self._lib.bh_linear_tetmesh_dq2(self._h, model, _ptr(q), _ptr(params), pp, _ptr(vals)),
                  "bh_linear_tetmesh_dq2")
        return vals

    def dq2(self, model, q, params, callback=None, model_ids=None):
        import scipy.sparse as sp

        rowptr, colind = self.pattern()
        vals = self.dq2_values(model, q, params, callback, model_ids)
        n = 3 * self.nV
        if callback is not None and not isinstance(callback, simple_psd_fix):
            # a host callback may leave the element matrices unsymmetric: convert the CSR values exactly
            return sp.csr_matrix((vals, colind, rowptr), shape=(n, n)).tocsc()
        # the pattern is symmetric: CSR arrays read as CSC give the transpose of the values, i.e.
        # H itself up to the rounding asymmetry of the element blocks
        return sp.csc_matrix((vals, colind, rowptr), shape=(n, n))

    # ---- device arrays (torch tensors on cuda:0; asynchronous on the given / current stream) --------
    @staticmethod
    def _stream(stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    @staticmethod
    def _dp(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None

    def q_device(self, model, q, params, out, model_ids=None, stream=None):
        check(self._lib.bh_linear_tetmesh_q_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                   self._dp(params), self._dp(out), self._stream(stream)),
              "bh_linear_tetmesh_q_device")

    def dq_device(self, model, q, params, out, model_ids=None, stream=None):
        check(self._lib.bh_linear_tetmesh_dq_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                    self._dp(params), self._dp(out), self._stream(stream)),
              "bh_linear_tetmesh_dq_device")

    def dq2_device(self, model, q, params, out, callback=None, model_ids=None, stream=None):
        psd = _psd_struct(callback)
        pp = ctypes.byref(psd) if psd is not None else None
        check(self._lib.bh_linear_tetmesh_dq2_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                     self._dp(params), pp, self._dp(out), self._stream(stream)),
              "bh_linear_tetmesh_dq2_device")

    # precision form: FP32 params (nT x 2 column-major, flattened) and FP32 Hessian values;
    # energy and gradient come back in FP64
    def q_f32_device(self, model, q, params32, out, model_ids=None, stream=None):
        check(self._lib.bh_linear_tetmesh_q_f32_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                       self._dp(params32), self._dp(out), self._stream(stream)),
              "bh_linear_tetmesh_q_f32_device")

    def dq_f32_device(self, model, q, params32, out, model_ids=None, stream=None):
        check(self._lib.bh_linear_tetmesh_dq_f32_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                        self._dp(params32), self._dp(out), self._stream(stream)),
              "bh_linear_tetmesh_dq_f32_device")

    def dq2_f32_device(self, model, q, params32, out, callback=None, model_ids=None, stream=None):
        psd = _psd_struct(callback)
        pp = ctypes.byref(psd) if psd is not None else None
        check(self._lib.bh_linear_tetmesh_dq2_f32_device(self._h, model, self._dp(model_ids), self._dp(q),
                                                         self._dp(params32), pp, self._dp(out),
                                                         self._stream(stream)), "bh_linear_tetmesh_dq2_f32_device")

    def mass_matrix_device(self, density, out, stream=None):
        check(self._lib.bh_linear_tetmesh_mass_matrix_device(self._h, self._dp(density), self._dp(out),
                                                             self._stream(stream)), "mass_matrix_device")

    def const_ext_force_device(self, traction, out, stream=None):
        t = (ctypes.c_double * 3)(*[float(x) for x in traction])
        check(self._lib.bh_linear_tetmesh_const_ext_force_device(self._h, t, self._dp(out), self._stream(stream)),
              "const_ext_force_device")

    def spmv_device(self, vals, x, y, stream=None):
        check(self._lib.bh_spmv_device(self._h, self._dp(vals), self._dp(x), self._dp(y), self._stream(stream)),
              "bh_spmv_device")

    def dphi_dX_device(self, V_colmajor, dphidX_out, vol_out, stream=None):
        check(self._lib.bh_linear_tetmesh_dphi_dX_device(self._h, self._dp(V_colmajor), self._dp(dphidX_out),
                                                         self._dp(vol_out), self._stream(stream)), "dphi_dX_device")


# ---- reference-shaped free functions ------------------------------------------------------------
_cache = {}


def _mesh_for(V, E, dphidX, volume):
    E = np.asarray(E, dtype=np.int32)
    dphidX = np.asarray(dphidX, dtype=np.float64)
    volume = np.asarray(volume, dtype=np.float64)
    h = hashlib.blake2b(digest_size=16)
    for a in (E, dphidX, volume):
        h.update(np.ascontiguousarray(a).tobytes())
    key = (int(np.asarray(V).shape[0]), h.hexdigest())
    m = _cache.get(key)
    if m is None:
        if len(_cache) > 8:
            _cache.clear()
        m = TetMesh(np.asarray(V).shape[0], E, dphidX, volume)
        _cache[key] = m
    return m


def _make(model):
    def q_fn(V, E, q, dphidX, volume, params):
        return _mesh_for(V, E, dphidX, volume).q(model, q, params)

    def dq_fn(V, E, q, dphidX, volume, params, callback=None):
        if callback is not None:
            raise NotImplementedError("gradient element callbacks are not supported on the GPU path")
        return _mesh_for(V, E, dphidX, volume).dq(model, q, params)

    def dq2_fn(V, E, q, dphidX, volume, params, callback=None):
        return _mesh_for(V, E, dphidX, volume).dq2(model, q, params, callback)

    return q_fn, dq_fn, dq2_fn


# src/linear_tetmesh_neohookean_{q,dq,dq2}.cpp
linear_tetmesh_neohookean_q, linear_tetmesh_neohookean_dq, linear_tetmesh_neohookean_dq2 = _make(NEOHOOKEAN)
# src/linear_tetmesh_stvk_{q,dq,dq2}.cpp
linear_tetmesh_stvk_q, linear_tetmesh_stvk_dq, linear_tetmesh_stvk_dq2 = _make(STVK)
# src/linear_tetmesh_corotational_{q,dq,dq2}.cpp
linear_tetmesh_corotational_q, linear_tetmesh_corotational_dq, linear_tetmesh_corotational_dq2 = _make(COROTATIONAL)
# src/linear_tetmesh_arap_{q,dq,dq2}.cpp
linear_tetmesh_arap_q, linear_tetmesh_arap_dq, linear_tetmesh_arap_dq2 = _make(ARAP)


def linear_tetmesh_generic_elastic_q(V, E, q, model_ids, dphidX, volume, params):
    """src/linear_tetmesh_generic_elastic_q.cpp:5-30 with the functor replaced by per-tet model ids."""
    return _mesh_for(V, E, dphidX, volume).q(MODEL_PER_TET, q, params, model_ids=model_ids)


def linear_tetmesh_generic_elastic_dq(V, E, q, model_ids, dphidX, volume, params):
    return _mesh_for(V, E, dphidX, volume).dq(MODEL_PER_TET, q, params, model_ids=model_ids)


def linear_tetmesh_generic_elastic_dq2(V, E, q, model_ids, dphidX, volume, params, callback=None):
    return _mesh_for(V, E, dphidX, volume).dq2(MODEL_PER_TET, q, params, callback, model_ids=model_ids)


# matlab/linear_tetmesh_neohookean_muscle_{q,dq,dq2}.cpp: Neo-Hookean (params_nh nT x 2) plus a muscle fibre
# per tet (directions nT x 3, params_a nT x 1); the element callback of dq2 as above
def muscle_params(directions, params_nh, params_a):
    """the nT x 6 parameter array of NEOHOOKEAN_MUSCLE: [C, D, a, d0, d1, d2]"""
    nh = np.asarray(params_nh, dtype=np.float64).reshape(-1, 2)
    a = np.asarray(params_a, dtype=np.float64).reshape(-1, 1)
    d = np.asarray(directions, dtype=np.float64).reshape(-1, 3)
    return np.asfortranarray(np.hstack([nh, a[:, :1], d]))


def linear_tetmesh_neohookean_muscle_q(V, E, q, dphidX, volume, directions, params_nh, params_a):
    return _mesh_for(V, E, dphidX, volume).q(NEOHOOKEAN_MUSCLE, q, muscle_params(directions, params_nh, params_a))


def linear_tetmesh_neohookean_muscle_dq(V, E, q, dphidX, volume, directions, params_nh, params_a):
    return _mesh_for(V, E, dphidX, volume).dq(NEOHOOKEAN_MUSCLE, q, muscle_params(directions, params_nh, params_a))


def linear_tetmesh_neohookean_muscle_dq2(V, E, q, dphidX, volume, directions, params_nh, params_a, callback=None):
    return _mesh_for(V, E, dphidX, volume).dq2(NEOHOOKEAN_MUSCLE, q, muscle_params(directions, params_nh, params_a),
                                               callback)

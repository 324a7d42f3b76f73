"""ctypes wrapper of the FP64 CPU oracle (oracle/oracle.cpp -> oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


class _CSC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("colptr", ctypes.POINTER(ctypes.c_int32)),
                ("rowind", ctypes.POINTER(ctypes.c_int32)), ("vals", ctypes.POINTER(ctypes.c_double))]


ENERGY_CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib_fast = None


def host_cpu():
    """(model name, logical CPUs usable by this process, a short tag of the model and its flags)"""
    import hashlib

    model, flags = "unknown", ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count() or 1
    return model, ncpu, hashlib.sha1((model + flags).encode()).hexdigest()[:10]


def fast_lib_path():
    return os.path.join(HERE, f"liboracle_fast_{host_cpu()[2]}.so")


def load(fast=False):
    """the checker build (-O2), or with fast=True the CPU-baseline timing build of the same source
    (-O3 -march=native; built on the host that runs it, since -march=native is host-specific)"""
    global _lib, _lib_fast
    if fast:
        if _lib_fast is None:
            path = fast_lib_path()
            # make rebuilds it when oracle.cpp is newer than a copy left by an earlier build
            subprocess.run(["make", "-s", "-C", HERE, "fast", "FAST_LIB=" + os.path.basename(path)], check=True)
            _lib_fast = _declare(ctypes.CDLL(path))
        return _lib_fast
    if _lib is not None:
        return _lib
    if os.environ.get("BH_ORACLE_SANITIZE") == "1":  # the AddressSanitizer / UBSan build (host checks)
        subprocess.run(["make", "-s", "-C", HERE, "sanitize"], check=True)
        _lib = _declare(ctypes.CDLL(os.path.join(HERE, "liboracle_san.so")))
        return _lib
    if not os.path.exists(LIB):
        build()
    _lib = _declare(ctypes.CDLL(LIB))
    return _lib


def _declare(L):
    vp, i, d, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
    L.or_dphidX_volume.argtypes = [i, vp, i, vp, vp, vp]
    L.or_dphidX_volume.restype = i
    L.or_psi.argtypes = [i, vp, vp]
    L.or_psi.restype = d
    L.or_dpsi.argtypes = [i, vp, vp, vp]
    L.or_d2psi.argtypes = [i, vp, vp, i, vp]
    L.or_svd3.argtypes = [vp, vp, vp, vp]
    L.or_psd_fix.argtypes = [vp, i, d]
    L.or_element_hessian.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, i, d, vp]
    L.or_tetmesh_q.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp]
    L.or_tetmesh_q.restype = d
    L.or_tetmesh_dq.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp]
    L.or_tetmesh_dq2.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, i, i, d]
    L.or_tetmesh_dq2.restype = ctypes.POINTER(_CSC)
    L.or_tetmesh_mass_matrix.argtypes = [i, i, vp, vp, vp]
    L.or_tetmesh_mass_matrix.restype = ctypes.POINTER(_CSC)
    L.or_csc_free.argtypes = [ctypes.POINTER(_CSC)]
    L.or_tetmesh_const_ext_force.argtypes = [i, i, vp, vp, vp, vp]
    L.or_newton_static.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, i, d, d, i, i, d, d, vp, vp,
                                   ctypes.POINTER(ctypes.c_int)]
    L.or_newton_static.restype = i
    L.or_newton_be.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, i, d, d, i, i, d, d, vp, vp,
                               ctypes.POINTER(ctypes.c_int)]
    L.or_newton_be.restype = i
    L.or_linesearch_backtracking.argtypes = [vp, vp, vp, i, ENERGY_CB, vp, i, d, d, d, ctypes.POINTER(ctypes.c_double)]
    L.or_linesearch_backtracking.restype = i
    L.or_band_solve.argtypes = [i64, vp, vp, vp, vp, vp]
    L.or_band_solve.restype = i
    L.or_psd_fix_ql.argtypes = [vp, d]
    L.or_tetmesh_dq2_timed.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, d, i, vp]
    L.or_tetmesh_dq2_timed.restype = ctypes.POINTER(_CSC)
    L.or_tri2d_dphidX.argtypes = [i, vp, i, vp, vp]
    L.or_tri2d_dphidX.restype = i
    L.or_tri3d_dphidX.argtypes = [i, vp, i, vp, vp]
    L.or_tri3d_dphidX.restype = i
    L.or_tri2dmesh_q.argtypes = [i, i, i, vp, vp, vp, vp, vp]
    L.or_tri2dmesh_q.restype = d
    L.or_tri2dmesh_dq.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp]
    L.or_tri2d_element_hessian.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, d, vp]
    L.or_tri2dmesh_dq2.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, d]
    L.or_tri2dmesh_dq2.restype = ctypes.POINTER(_CSC)
    L.or_tri2dmesh_mass.argtypes = [i, i, vp, vp, vp]
    L.or_tri2dmesh_mass.restype = ctypes.POINTER(_CSC)
    L.or_tri3d_stvk_element_gradient.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_tri3d_stvk_element_hessian.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, i, d, vp]
    L.or_tri3dmesh_stvk_dq2.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, i, d]
    L.or_tri3dmesh_stvk_dq2.restype = ctypes.POINTER(_CSC)
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _F(a, dt):
    return np.asfortranarray(np.asarray(a, dtype=dt))


def _take_csc(ptr):
    import scipy.sparse as sp

    c = ptr.contents
    n, nnz = c.n, c.nnz
    colptr = np.ctypeslib.as_array(c.colptr, shape=(n + 1,)).copy()
    rowind = np.ctypeslib.as_array(c.rowind, shape=(max(nnz, 1),))[:nnz].copy()
    vals = np.ctypeslib.as_array(c.vals, shape=(max(nnz, 1),))[:nnz].copy()
    load().or_csc_free(ptr)
    return sp.csc_matrix((vals, rowind, colptr), shape=(n, n))


def dphidX_volume(V, E):
    L = load()
    V = _F(V, np.float64)
    E = _F(E, np.int32)
    dX = np.zeros((E.shape[0], 12), order="F")
    vol = np.zeros(E.shape[0])
    assert L.or_dphidX_volume(V.shape[0], _p(V), E.shape[0], _p(E), _p(dX), _p(vol)) == 0
    return dX, vol


def psi(model, F, p):
    F = np.asarray(F, dtype=np.float64).reshape(3, 3)
    f = np.asfortranarray(F)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    return load().or_psi(model, _p(f), _p(pp))


def dpsi(model, F, p):
    f = np.asfortranarray(np.asarray(F, dtype=np.float64).reshape(3, 3))
    pp = np.ascontiguousarray(p, dtype=np.float64)
    P = np.zeros(9)
    load().or_dpsi(model, _p(f), _p(pp), _p(P))
    return P.reshape(3, 3, order="F")


def d2psi(model, F, p, ref_layout=1):
    """9x9 in vec(F) column-major ordering."""
    f = np.asfortranarray(np.asarray(F, dtype=np.float64).reshape(3, 3))
    pp = np.ascontiguousarray(p, dtype=np.float64)
    H = np.zeros(81)
    load().or_d2psi(model, _p(f), _p(pp), ref_layout, _p(H))
    return H.reshape(9, 9, order="F")


def svd3(F):
    f = np.asfortranarray(np.asarray(F, dtype=np.float64).reshape(3, 3))
    U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
    load().or_svd3(_p(f), _p(U), _p(S), _p(V))
    return U.reshape(3, 3, order="F"), S, V.reshape(3, 3, order="F")


def psd_fix_ql(A, tol):
    """12x12 simple_psd_fix by tridiagonalisation + implicit QR (the timing build's eigensolver)"""
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    assert A.shape == (12, 12)
    load().or_psd_fix_ql(_p(A), tol)
    return A


def tetmesh_dq2_timed(model, nV, E, q, dX, vol, params, psd=False, psd_tol=1e-3, threads=1, fast=True):
    """the CPU baseline's timing mode of tetmesh_dq2: (H, [element + triplet seconds, setFromTriplets
    seconds]) with the elements split over `threads` OpenMP threads"""
    L = load(fast=fast)
    E, dX, params = _F(E, np.int32), _F(dX, np.float64), _F(params, np.float64)
    q, vol = np.ascontiguousarray(q, dtype=np.float64), np.ascontiguousarray(vol, dtype=np.float64)
    times = np.zeros(2)
    ptr = L.or_tetmesh_dq2_timed(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params), 1 if psd else 0,
                                 psd_tol, threads, _p(times))
    return _take_csc(ptr), times


def psd_fix(A, tol):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    load().or_psd_fix(_p(A), A.shape[0], tol)
    return A


def element_hessian(model, e, E, q, dX, vol, params, ref_layout=1, psd=False, psd_tol=1e-3):
    E, dX, params = _F(E, np.int32), _F(dX, np.float64), _F(params, np.float64)
    q, vol = np.ascontiguousarray(q, dtype=np.float64), np.ascontiguousarray(vol, dtype=np.float64)
    K = np.zeros(144)
    load().or_element_hessian(model, E.shape[0], e, _p(E), _p(q), _p(dX), _p(vol), _p(params), ref_layout,
                              1 if psd else 0, psd_tol, _p(K))
    return K.reshape(12, 12, order="F")


def _args(E, q, dX, vol, params, model_ids):
    E, dX, params = _F(E, np.int32), _F(dX, np.float64), _F(params, np.float64)
    q, vol = np.ascontiguousarray(q, dtype=np.float64), np.ascontiguousarray(vol, dtype=np.float64)
    ids = None if model_ids is None else np.ascontiguousarray(model_ids, dtype=np.uint8)
    return E, q, dX, vol, params, ids


def tetmesh_q(model, nV, E, q, dX, vol, params, model_ids=None):
    E, q, dX, vol, params, ids = _args(E, q, dX, vol, params, model_ids)
    return load().or_tetmesh_q(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params), _p(ids))


def tetmesh_dq(model, nV, E, q, dX, vol, params, model_ids=None):
    E, q, dX, vol, params, ids = _args(E, q, dX, vol, params, model_ids)
    g = np.zeros(3 * nV)
    load().or_tetmesh_dq(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params), _p(ids), _p(g))
    return g


def tetmesh_dq2(model, nV, E, q, dX, vol, params, model_ids=None, ref_layout=1, psd=False, psd_tol=1e-3):
    """scipy CSC with the reference's pattern and fold order."""
    E, q, dX, vol, params, ids = _args(E, q, dX, vol, params, model_ids)
    ptr = load().or_tetmesh_dq2(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params), _p(ids),
                                ref_layout, 1 if psd else 0, psd_tol)
    return _take_csc(ptr)


def mass_matrix(nV, E, density, vol):
    E = _F(E, np.int32)
    density = np.ascontiguousarray(np.broadcast_to(density, (E.shape[0],)), dtype=np.float64)
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    return _take_csc(load().or_tetmesh_mass_matrix(nV, E.shape[0], _p(E), _p(density), _p(vol)))


def const_ext_force(nV, E, vol, traction):
    E = _F(E, np.int32)
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    t = np.ascontiguousarray(traction, dtype=np.float64)
    f = np.zeros(3 * nV)
    load().or_tetmesh_const_ext_force(nV, E.shape[0], _p(E), _p(vol), _p(t), _p(f))
    return f


def newton_static(model, nV, E, dX, vol, params, q0, fixed_dof=None, f_ext=None, model_ids=None, psd=False,
                  psd_tol=1e-3, tol=1e-3, max_it=100, max_ls=100, c=1e-8, p=0.5):
    """Reference-semantics Newton with the banded direct solve.  Returns (q, iterations, converged, trace)."""
    E, q, dX, vol, params, ids = _args(E, q0, dX, vol, params, model_ids)
    q = q.copy()
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, dtype=np.uint8)
    fe = None if f_ext is None else np.ascontiguousarray(f_ext, dtype=np.float64)
    trace = np.zeros(4 * (max_it + 1))
    conv = ctypes.c_int()
    it = load().or_newton_static(model, nV, E.shape[0], _p(E), _p(dX), _p(vol), _p(params), _p(ids), _p(fx), _p(fe),
                                 1 if psd else 0, psd_tol, tol, max_it, max_ls, c, p, _p(q), _p(trace),
                                 ctypes.byref(conv))
    if it < 0:
        raise RuntimeError("oracle direct solve failed")
    return q, it, bool(conv.value), trace.reshape(-1, 4)[:it + 1]



def newton_be(model, nV, E, dX, vol, params, q0, q_pred, dt, density, fixed_dof=None, f_ext=None, model_ids=None,
              psd=False, psd_tol=1e-3, tol=1e-3, max_it=100, max_ls=100, c=1e-8, p=0.5):
    """One backward-Euler step (position form of matlab/examples/test_tet_materials.m:97-112) with the
    reference Newton semantics and the banded direct solve.  density: per tet.  Returns
    (q, iterations, converged, trace)."""
    E, q, dX, vol, params, ids = _args(E, q0, dX, vol, params, model_ids)
    q = q.copy()
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, dtype=np.uint8)
    fe = None if f_ext is None else np.ascontiguousarray(f_ext, dtype=np.float64)
    rho = np.ascontiguousarray(np.broadcast_to(np.asarray(density, dtype=np.float64), (E.shape[0],)))
    qp = np.ascontiguousarray(q_pred, dtype=np.float64)
    trace = np.zeros(4 * (max_it + 1))
    conv = ctypes.c_int()
    it = load().or_newton_be(model, nV, E.shape[0], _p(E), _p(dX), _p(vol), _p(params), _p(ids), _p(fx), _p(fe),
                             _p(rho), _p(qp), dt, 1 if psd else 0, psd_tol, tol, max_it, max_ls, c, p, _p(q),
                             _p(trace), ctypes.byref(conv))
    if it < 0:
        raise RuntimeError("oracle direct solve failed")
    return q, it, bool(conv.value), trace.reshape(-1, 4)[:it + 1]

def linesearch(x, d, g, f, max_it=100, alpha=1.0, c=1e-8, p=0.5):
    """reference linesearch_backtracking_bisection on a Python energy f(x)."""
    x = np.array(x, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    n = x.size

    @ENERGY_CB
    def cb(ptr, nn, user):
        return float(f(np.ctypeslib.as_array(ptr, shape=(nn,)).copy()))

    a = ctypes.c_double()
    st = load().or_linesearch_backtracking(_p(x), _p(d), _p(g), n, cb, None, max_it, alpha, c, p, ctypes.byref(a))
    return x, a.value, st


def band_solve(A, b):
    import scipy.sparse as sp

    A = sp.csc_matrix(A)
    A.sort_indices()
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    rowind = np.ascontiguousarray(A.indices, dtype=np.int32)
    vals = np.ascontiguousarray(A.data, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    if load().or_band_solve(A.shape[0], _p(colptr), _p(rowind), _p(vals), _p(b), _p(x)) != 0:
        raise RuntimeError("band solve failed")
    return x


# ---- triangle siblings (tri2d plane-strain elements, tri3d StVK shell) -----------------------------
def tri2d_dphidX(V, E):
    """nT x 6 (Phi(m, j) at column m + 3j) of 2-D triangles (src/linear_tri2dmesh_dphi_dX.cpp)"""
    V, E = _F(V, np.float64), _F(E, np.int32)
    out = np.zeros((E.shape[0], 6), order="F")
    assert load().or_tri2d_dphidX(V.shape[0], _p(V), E.shape[0], _p(E), _p(out)) == 0
    return out


def tri3d_dphidX(V, E):
    """nT x 9 (Phi(m, j) at column m + 3j) of triangles in 3-D (src/linear_tri3dmesh_dphi_dX.cpp)"""
    V, E = _F(V, np.float64), _F(E, np.int32)
    out = np.zeros((E.shape[0], 9), order="F")
    assert load().or_tri3d_dphidX(V.shape[0], _p(V), E.shape[0], _p(E), _p(out)) == 0
    return out


def _tri_args(E, q, dX, vol, params):
    return (_F(E, np.int32), np.ascontiguousarray(q, dtype=np.float64), _F(dX, np.float64),
            np.ascontiguousarray(vol, dtype=np.float64), _F(params, np.float64))


def tri2dmesh_q(model, nV, E, q, dX, vol, params):
    E, q, dX, vol, params = _tri_args(E, q, dX, vol, params)
    return load().or_tri2dmesh_q(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params))


def tri2dmesh_dq(model, nV, E, q, dX, vol, params):
    E, q, dX, vol, params = _tri_args(E, q, dX, vol, params)
    g = np.zeros(2 * nV)
    load().or_tri2dmesh_dq(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params), _p(g))
    return g


def tri2d_element_hessian(model, e, E, q, dX, vol, params, psd=False, psd_tol=1e-3):
    E, q, dX, vol, params = _tri_args(E, q, dX, vol, params)
    K = np.zeros((6, 6), order="F")
    load().or_tri2d_element_hessian(model, E.shape[0], e, _p(E), _p(q), _p(dX), _p(vol), _p(params),
                                    1 if psd else 0, psd_tol, _p(K))
    return K


def tri2dmesh_dq2(model, nV, E, q, dX, vol, params, psd=False, psd_tol=1e-3):
    E, q, dX, vol, params = _tri_args(E, q, dX, vol, params)
    return _take_csc(load().or_tri2dmesh_dq2(model, nV, E.shape[0], _p(E), _p(q), _p(dX), _p(vol), _p(params),
                                             1 if psd else 0, psd_tol))


def tri2dmesh_mass(nV, E, density, vol):
    E = _F(E, np.int32)
    density, vol = np.ascontiguousarray(density, dtype=np.float64), np.ascontiguousarray(vol, dtype=np.float64)
    return _take_csc(load().or_tri2dmesh_mass(nV, E.shape[0], _p(E), _p(density), _p(vol)))


def _tri3_args(E, q, dX, N, n, dndq, vol, params):
    E, q, dX, vol, params = _tri_args(E, q, dX, vol, params)
    return E, q, dX, _F(N, np.float64), _F(n, np.float64), _F(dndq, np.float64), vol, params


def tri3d_stvk_element_gradient(e, E, q, dX, N, n, dndq, vol, params):
    E, q, dX, N, n, dndq, vol, params = _tri3_args(E, q, dX, N, n, dndq, vol, params)
    g = np.zeros(9)
    load().or_tri3d_stvk_element_gradient(E.shape[0], e, _p(E), _p(q), _p(dX), _p(N), _p(n), _p(dndq), _p(vol),
                                          _p(params), _p(g))
    return g


def tri3d_stvk_element_hessian(e, E, q, dX, N, n, dndq, vol, params, psd=False, psd_tol=1e-3):
    E, q, dX, N, n, dndq, vol, params = _tri3_args(E, q, dX, N, n, dndq, vol, params)
    K = np.zeros((9, 9), order="F")
    load().or_tri3d_stvk_element_hessian(E.shape[0], e, _p(E), _p(q), _p(dX), _p(N), _p(n), _p(dndq), _p(vol),
                                         _p(params), 1 if psd else 0, psd_tol, _p(K))
    return K


def tri3dmesh_stvk_dq2(nV, E, q, dX, N, n, dndq, vol, params, psd=False, psd_tol=1e-3):
    E, q, dX, N, n, dndq, vol, params = _tri3_args(E, q, dX, N, n, dndq, vol, params)
    return _take_csc(load().or_tri3dmesh_stvk_dq2(nV, E.shape[0], _p(E), _p(q), _p(dX), _p(N), _p(n), _p(dndq),
                                                  _p(vol), _p(params), 1 if psd else 0, psd_tol))

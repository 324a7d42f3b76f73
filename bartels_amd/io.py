"""Host-side file formats over the C ABI (no GPU call): TetGen and MEDIT meshes, and MatrixMarket
sparse matrices in the exact format of Eigen's saveMarket / loadMarket.

Names follow the reference: ``read_tetgen(node_file, ele_file) -> (V, E)`` (src/read_tetgen.cpp:105-124),
``read_mesh(path) -> (V, E)`` for MEDIT (igl::readMESH as tests/test_assemble.cpp:25 uses it),
``save_market(A, path)`` / ``load_market(path)`` (Eigen::saveMarket / loadMarket; A a scipy CSC matrix,
its arrays exactly Eigen::SparseMatrix's).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _enc(path):
    return str(path).encode()


def _mesh(call, *paths):
    lib = _lib.load()
    nV, nT = ctypes.c_int(), ctypes.c_int()
    check(call(lib)(*[_enc(p) for p in paths], ctypes.byref(nV), ctypes.byref(nT), None, None), "read mesh")
    V = np.zeros((nV.value, 3), order="F")
    E = np.zeros((nT.value, 4), dtype=np.int32, order="F")
    check(call(lib)(*[_enc(p) for p in paths], ctypes.byref(nV), ctypes.byref(nT), V.ctypes.data_as(ctypes.c_void_p),
                    E.ctypes.data_as(ctypes.c_void_p)), "read mesh")
    return V, E


def read_tetgen(node_file, ele_file):
    """(V nV x 3, E nT x 4) as written in the files (indices not renumbered, as the reference reads them)"""
    return _mesh(lambda lib: lib.bh_read_tetgen, node_file, ele_file)


def read_mesh(path):
    """MEDIT .mesh: (V nV x 3, E nT x 4, 0-based)"""
    return _mesh(lambda lib: lib.bh_read_medit, path)


def save_market(A, path):
    import scipy.sparse as sp

    A = sp.csc_matrix(A)
    A.sort_indices()
    cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ri = np.ascontiguousarray(A.indices, dtype=np.int32)
    vv = np.ascontiguousarray(A.data, dtype=np.float64)
    check(_lib.load().bh_save_market(_enc(path), A.shape[0], A.shape[1], cp.ctypes.data_as(ctypes.c_void_p),
                                     ri.ctypes.data_as(ctypes.c_void_p), vv.ctypes.data_as(ctypes.c_void_p)),
          "bh_save_market")


def load_market(path):
    import scipy.sparse as sp

    lib = _lib.load()
    m, n, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib.bh_load_market(_enc(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz), None, None, None),
          "bh_load_market")
    cp = np.zeros(n.value + 1, np.int32)
    ri = np.zeros(nnz.value, np.int32)
    vv = np.zeros(nnz.value)
    check(lib.bh_load_market(_enc(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                             cp.ctypes.data_as(ctypes.c_void_p), ri.ctypes.data_as(ctypes.c_void_p),
                             vv.ctypes.data_as(ctypes.c_void_p)), "bh_load_market")
    return sp.csc_matrix((vv, ri, cp), shape=(m.value, n.value))

"""The arbitrary element-matrix callback (bh_linear_tetmesh_dq2_callback, the reference's
ElementMatrixCallback: src/linear_tetmesh_neohookean_dq2.cpp:17-23, include/EigenTypes.h:133).

- The matrix the callback receives is the reference's element matrix: the oracle's 12x12 of the same
  tet (tests/test_oracle.py pins that against autograd and the bunny fixtures), same layout, to 1e-12.
- A callback that changes nothing assembles the values of bh_linear_tetmesh_dq2 without PSD to
  rounding (the same closed-form blocks; the two kernels may fuse multiply-adds differently), for
  every model.
- Past the 262,144-tet chunk size, where later chunks add onto the values of earlier ones: a callback
  that writes small integers assembles exactly (integer sums are exact in FP64) what numpy's scatter
  of the same matrices gives, bit for bit.
- A callback that projects each element matrix with numpy's eigh (the spectral definition of
  simple_psd_fix, src/simple_psd_fix.cpp:5-21) agrees with the device PSD path and the oracle.
- A callback that makes the element matrices unsymmetric (a different rank-one term per tet)
  assembles exactly what numpy's scatter of those matrices gives.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


def nrel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def _problem(grid, seed, model):
    from bartels_amd import tetmesh as tm

    nx, ny, nz = grid
    V, E = meshgen.beam_grid(nx, ny, nz, 0.05)
    V = meshgen.jitter(V, E, 0.05, 0.1, seed)
    nV, nT = V.shape[0], E.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    rng = np.random.default_rng(seed)
    q = V.ravel() + 0.002 * rng.standard_normal(3 * nV)
    lam, mu = meshgen.emu_to_lame(np.exp(rng.uniform(np.log(1e5), np.log(1e7), nT)), 0.45)
    ids = np.full(nT, min(model, 4), np.uint8)
    params = meshgen.model_params(ids, lam, mu)
    if model == tm.NEOHOOKEAN_MUSCLE:
        d = rng.standard_normal((nT, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        params = tm.muscle_params(d, params, rng.uniform(1e4, 1e5, nT))
    return V, E, q, dX, vol, np.asfortranarray(params)


def _vals(mesh, model, q, params, callback=None, model_ids=None):
    return mesh.dq2_values(model, q, params, callback, model_ids)


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4, 5])
def test_callback_sees_reference_element_matrix_and_identity_matches(oracle, model):
    from bartels_amd import tetmesh as tm

    V, E, q, dX, vol, params = _problem((8, 4, 4), 11 + model, model)
    mesh = tm.TetMesh(V.shape[0], E, dX, vol, V=V)
    try:
        seen = []
        worst = [0.0]

        def check(H):
            e = len(seen)
            seen.append(e)
            if e % 7 == 0 and model != tm.NEOHOOKEAN_MUSCLE:
                K = oracle.element_hessian(model, e, E, q, dX, vol, params)
                worst[0] = max(worst[0], nrel(H, K))

        plain = _vals(mesh, model, q, params)
        got = _vals(mesh, model, q, params, callback=check)
        diff = nrel(got, plain)
        print(f"model {model}: {len(seen)} callbacks, element matrices vs oracle {worst[0]:.2e}, "
              f"identity callback vs dq2 {diff:.2e} (bit-exact: {np.array_equal(got, plain)})")
        assert seen == list(range(E.shape[0]))  # once per tet, ascending element order
        assert worst[0] < 1e-12
        assert diff < 1e-14
    finally:
        mesh.close()


def test_callback_psd_matches_device_psd_and_oracle(oracle):
    from bartels_amd import tetmesh as tm

    V, E, q, dX, vol, params = _problem((8, 4, 4), 5, 0)
    # a larger deformation, so that the projection changes many element matrices
    q = V.ravel() + 0.01 * np.random.default_rng(6).standard_normal(V.size)
    mesh = tm.TetMesh(V.shape[0], E, dX, vol, V=V)
    tol = 1e-3

    def psd(H):
        w, U = np.linalg.eigh(0.5 * (H + H.T))
        H[:] = (U * np.maximum(w, tol)) @ U.T

    try:
        got = tm.linear_tetmesh_neohookean_dq2(V, E, q, dX, vol, params, callback=psd)
        rowptr, colind = mesh.pattern()
        n = 3 * mesh.nV
        dev = sp.csr_matrix((_vals(mesh, 0, q, params, callback=tm.simple_psd_fix(tol)), colind, rowptr), shape=(n, n))
        ref = oracle.tetmesh_dq2(0, mesh.nV, E, q, dX, vol, params, psd=True, psd_tol=tol)
        e_dev = sp.linalg.norm(got - dev) / sp.linalg.norm(dev)
        e_or = sp.linalg.norm(got - ref) / sp.linalg.norm(ref)
        print(f"numpy-eigh callback vs device PSD {e_dev:.2e}, vs oracle {e_or:.2e}")
        assert e_dev < 1e-11 and e_or < 1e-11
    finally:
        mesh.close()


def test_callback_unsymmetric_and_per_tet_models():
    from bartels_amd import tetmesh as tm

    V, E, q, dX, vol, params = _problem((7, 4, 3), 9, 0)
    nV, nT = V.shape[0], E.shape[0]
    ids = (np.arange(nT) % 5).astype(np.uint8)
    rng = np.random.default_rng(3)
    U = rng.standard_normal((nT, 12))
    W = rng.standard_normal((nT, 12))
    mesh = tm.TetMesh(nV, E, dX, vol, V=V)
    try:
        count = [0]

        def skew(H):
            e = count[0]
            count[0] += 1
            H += 1e3 * np.outer(U[e], W[e])

        got = mesh.dq2(tm.MODEL_PER_TET, q, params, callback=skew, model_ids=ids)
        plain = mesh.dq2(tm.MODEL_PER_TET, q, params, model_ids=ids)
        rowptr, colind = mesh.pattern()
        n = 3 * nV
        plain_csr = sp.csr_matrix((mesh.dq2_values(tm.MODEL_PER_TET, q, params, model_ids=ids), colind, rowptr),
                                  shape=(n, n))
        dof = (3 * E[:, :, None] + np.arange(3)).reshape(nT, 12)  # row 3 a + c -> 3 E(e, a) + c
        rows = np.repeat(dof, 12, axis=1).ravel()
        cols = np.tile(dof, (1, 12)).ravel()
        add = 1e3 * (U[:, :, None] * W[:, None, :]).reshape(nT, 144)
        P = sp.csr_matrix((add.ravel(), (rows, cols)), shape=(n, n))
        want = plain_csr + P
        err = sp.linalg.norm(got - want) / sp.linalg.norm(want)
        asym = sp.linalg.norm(got - got.T) / sp.linalg.norm(got)
        print(f"unsymmetric callback, per-tet models: {count[0]} callbacks, vs numpy scatter {err:.2e}, "
              f"asymmetry of the result {asym:.2e}")
        assert count[0] == nT
        assert err < 1e-14
        assert asym > 1e-3
        assert sp.linalg.norm(plain - plain_csr.tocsc()) / sp.linalg.norm(plain_csr) < 1e-14
    finally:
        mesh.close()


def test_callback_across_chunks_is_exact():
    """304,842 tets: two chunks of the staged assembly (262,144 + 42,698).  The callback replaces each
    element matrix with integers in [-8, 8] drawn from the element index; every assembled value is
    an exact integer sum, so the result must equal numpy's scatter of the same matrices bit for bit
    (a tet assembled twice or not at all, or a chunk that overwrote instead of adding, shows)"""
    from bartels_amd import tetmesh as tm

    V, E, q, dX, vol, params = _problem((48, 24, 48), 2, 0)
    nV, nT = V.shape[0], E.shape[0]
    assert nT > 262144
    mesh = tm.TetMesh(nV, E, dX, vol, V=V)

    def ints(e):
        return ((np.arange(144) * 7 + 13 * e) % 17 - 8).astype(np.float64)

    try:
        n = [0]

        def write(H):
            H[:] = ints(n[0]).reshape(12, 12, order="F")
            n[0] += 1

        got = _vals(mesh, 0, q, params, callback=write)
        rowptr, colind = mesh.pattern()
        N = 3 * nV
        dof = (3 * E[:, :, None] + np.arange(3)).reshape(nT, 12)
        K = np.stack([ints(e) for e in range(nT)]).reshape(nT, 12, 12).transpose(0, 2, 1)  # [e, row, col]
        rows = np.repeat(dof, 12, axis=1).ravel()
        cols = np.tile(dof, (1, 12)).ravel()
        want = sp.csr_matrix((K.reshape(-1), (rows, cols)), shape=(N, N))
        want.sum_duplicates()
        gotm = sp.csr_matrix((got, colind, rowptr), shape=(N, N))
        d = abs(gotm - want)
        print(f"{nT} tets, {n[0]} callbacks, largest difference from the integer scatter {d.max() if d.nnz else 0.0}")
        assert n[0] == nT
        assert (d.nnz == 0) or d.max() == 0.0
    finally:
        mesh.close()


def test_callback_device_entry_point():
    """the device variant with a C-level callback (a ctypes function) on device arrays"""
    import torch

    from bartels_amd import _lib
    from bartels_amd import tetmesh as tm

    V, E, q, dX, vol, params = _problem((6, 3, 3), 4, 0)
    mesh = tm.TetMesh(V.shape[0], E, dX, vol, V=V)
    try:
        dq = torch.tensor(q, device="cuda")
        dp = torch.tensor(params.ravel(order="F"), device="cuda")
        out = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
        ref = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")

        def double_it(ptr, element, user):
            for i in range(144):
                ptr[i] *= 2.0

        cb = _lib.ELEMENT_CALLBACK(double_it)
        lib = _lib.load()
        assert lib.bh_linear_tetmesh_dq2_callback_device(mesh.handle, 0, None, ctypes.c_void_p(dq.data_ptr()),
                                                         ctypes.c_void_p(dp.data_ptr()), cb, None,
                                                         ctypes.c_void_p(out.data_ptr()), None) == 0
        ident = _lib.ELEMENT_CALLBACK(lambda ptr, element, user: None)
        assert lib.bh_linear_tetmesh_dq2_callback_device(mesh.handle, 0, None, ctypes.c_void_p(dq.data_ptr()),
                                                         ctypes.c_void_p(dp.data_ptr()), ident, None,
                                                         ctypes.c_void_p(ref.data_ptr()), None) == 0
        plain = torch.empty_like(ref)
        mesh.dq2_device(0, dq, dp, plain)
        torch.cuda.synchronize()
        print(f"doubling callback: max |out - 2 ref| = {float((out - 2 * ref).abs().max()):.2e}, identity vs "
              f"dq2 {float(torch.linalg.norm(ref - plain) / torch.linalg.norm(plain)):.2e}")
        assert torch.equal(out, 2 * ref)  # doubling is exact, and so is every sum of doubled terms
        assert float(torch.linalg.norm(ref - plain) / torch.linalg.norm(plain)) < 1e-14
        assert lib.bh_linear_tetmesh_dq2_callback_device(mesh.handle, 0, None, ctypes.c_void_p(dq.data_ptr()),
                                                         ctypes.c_void_p(dp.data_ptr()), _lib.ELEMENT_CALLBACK(),
                                                         None, ctypes.c_void_p(out.data_ptr()), None) == _lib.BH_BAD_ARG
    finally:
        mesh.close()

"""Pins the FP64 CPU oracle before it is trusted as the checker of the GPU path.

Each check is independent of the oracle's own formulas:
  * the reference's own fixtures (bunny mesh, mass matrix, legacy-NH Hessian, F = diag(2,4,3));
  * torch-FP64 autograd golden vectors of each energy (tests/golden/make_golden.py);
  * central finite differences of the oracle's own energy (h = 1e-6 as src/finite_differences.cpp:34-40);
  * invariants (rest state, frame and translation invariance, symmetry);
  * numpy eigh for the spectral definition of simple_psd_fix;
  * closed-form line-search and linear-solve cases.
"""
import hashlib
import os

import numpy as np
import pytest
import scipy.sparse as sp

from bartels_amd import meshgen
from conftest import REFERENCE

MODELS = [0, 1, 2, 3, 4]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---- reference fixtures -----------------------------------------------------------------------
def bunny_geometry(oracle, bunny):
    V, E = bunny["V"], bunny["E"]
    dX, vol = oracle.dphidX_volume(V, E)
    return V, E, dX, vol


def test_fixture_provenance(bunny):
    if not os.path.isdir(REFERENCE):
        pytest.skip("reference not mounted")
    files = ["data/meshes_mesh/coarser_bunny.mesh", "data/matrices/M_coarse_bunny.txt",
             "data/matrices/H_neohookean_coarser_bunny.txt"]
    for f, d in zip(files, bunny["sha256"]):
        assert hashlib.sha256(open(os.path.join(REFERENCE, f), "rb").read()).hexdigest() == d


def test_known_answer_F_diag(oracle, bunny):
    """tests/test_constitutive_model.cpp:42-66: X scaled by diag(2,4,3) gives F = diag(2,4,3) in every tet."""
    V, E, dX, vol = bunny_geometry(oracle, bunny)
    assert (vol > 0).all()
    q = (V * np.array([2.0, 4.0, 3.0])).ravel()
    for e in range(E.shape[0]):
        phi = dX[e].reshape(4, 3, order="F")
        X = q.reshape(-1, 3)[E[e]]  # 4 x 3 positions
        F = X.T @ phi
        assert np.abs(F - np.diag([2.0, 4.0, 3.0])).max() < 1e-8


def test_mass_matrix_fixture(oracle, bunny):
    """tests/test_assemble.cpp:25-46: M (density 0.1) equals M_coarse_bunny.txt; pattern bit-exact."""
    V, E, dX, vol = bunny_geometry(oracle, bunny)
    M = oracle.mass_matrix(V.shape[0], E, 0.1, vol)
    M.sort_indices()
    assert np.array_equal(M.indptr, bunny["indptr"]) and np.array_equal(M.indices, bunny["indices"])
    assert np.linalg.norm(M.data - bunny["M"]) < 1e-8  # the reference's criterion (absolute)
    assert np.linalg.norm(M.data - bunny["M"]) / bunny["M"][0] < 1e-12
    # explicit zeros kept: 48,312 stored entries, only a third numerically non-zero
    assert M.nnz == 48312 and np.count_nonzero(M.data) == 16104


def test_ext_force_is_M_g(oracle, bunny):
    """tests/test_assemble.cpp:54-69: const_ext_force(0.1 g) == M * g."""
    V, E, dX, vol = bunny_geometry(oracle, bunny)
    g = np.array([0.0, -9.8, 0.0])
    f = oracle.const_ext_force(V.shape[0], E, vol, 0.1 * g)
    M = oracle.mass_matrix(V.shape[0], E, 0.1, vol)
    assert np.linalg.norm(f - M @ np.tile(g, V.shape[0])) < 1e-8


def test_hessian_fixture_legacy_nh(oracle, bunny):
    """tests/test_constitutive_model.cpp:73-103: the H fixture pins the legacy isochoric NH."""
    V, E, dX, vol = bunny_geometry(oracle, bunny)
    q = (V * np.array([2.0, 4.0, 3.0])).ravel()
    YM, nu = 6e5, 0.4
    D = 0.5 * (YM * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    C = 0.5 * YM / (2.0 * (1.0 + nu))
    params = np.tile([C, D], (E.shape[0], 1))
    H = oracle.tetmesh_dq2(4, V.shape[0], E, q, dX, vol, params)
    H.sort_indices()
    assert np.array_equal(H.indptr, bunny["indptr"]) and np.array_equal(H.indices, bunny["indices"])
    err = np.linalg.norm(H.data - bunny["H"]) / H[0, 0]
    assert err < 1e-12, err
    # the active log-barrier form does not reproduce it (the reference's own test would print FAILED)
    Ha = oracle.tetmesh_dq2(0, V.shape[0], E, q, dX, vol, params)
    assert np.linalg.norm(Ha.data - bunny["H"]) / Ha[0, 0] > 0.1


# ---- autograd golden vectors ----------------------------------------------------------------
@pytest.mark.parametrize("model", MODELS)
def test_elements_vs_autograd(oracle, elements_golden, model):
    g = elements_golden
    F, p = g[f"F{model}"], g[f"p{model}"]
    for i in range(F.shape[0]):
        Fm = F[i].reshape(3, 3, order="F")
        assert abs(oracle.psi(model, Fm, p[i]) - g[f"psi{model}"][i]) <= 1e-12 * max(1.0, abs(g[f"psi{model}"][i]))
        assert rel(oracle.dpsi(model, Fm, p[i]).reshape(-1, order="F"), g[f"P{model}"][i]) < 1e-12
        Hr = oracle.d2psi(model, Fm, p[i], 1).reshape(-1, order="F")
        ref = g[f"Href{model}"][i] if model in (2, 3) else g[f"H{model}"][i]
        assert rel(Hr, ref) < 1e-12
        Ht = oracle.d2psi(model, Fm, p[i], 0).reshape(-1, order="F")
        assert rel(Ht, g[f"H{model}"][i]) < 1e-12


@pytest.mark.parametrize("model", MODELS)
def test_rest_state_and_frame_invariance(oracle, model):
    p = np.array([3e5, 7e5])
    I = np.eye(3)
    assert abs(oracle.psi(model, I, p)) < 1e-9
    assert np.abs(oracle.dpsi(model, I, p)).max() < 1e-9
    rng = np.random.default_rng(model)
    F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    if np.linalg.det(F) < 0:
        F[:, 0] *= -1
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    assert abs(oracle.psi(model, Q @ F, p) - oracle.psi(model, F, p)) < 1e-9 * max(1, abs(oracle.psi(model, F, p)))
    H = oracle.d2psi(model, F, p, 0)
    assert rel(H, H.T) < 1e-12


# ---- mesh level: finite differences and invariants -------------------------------------------
def muscle_columns(nT, lam, seed):
    """[a, d0, d1, d2] of the muscle fibre per tet: stiffness of the order of lambda, random unit directions"""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nT, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.hstack([(np.asarray(lam) * (0.5 + rng.random(nT)))[:, None], d])


def small_problem(model, seed=0, grid=(5, 3, 3), amp=0.08):
    V, E = meshgen.beam_grid(*grid, h=0.1)
    rng = np.random.default_rng(seed)
    q = (V + amp * 0.1 * rng.standard_normal(V.shape)).ravel()
    lam, mu = meshgen.emu_to_lame(1e5 * (1 + rng.random(E.shape[0])), 0.3 + 0.1 * rng.random(E.shape[0]))
    if model == 5:  # Neo-Hookean plus a muscle fibre: nT x 6 [C, D, a, d]
        nh = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
        return V, E, q, np.asfortranarray(np.hstack([nh, muscle_columns(E.shape[0], lam, seed + 1)]))
    ids = np.full(E.shape[0], model, dtype=np.uint8)
    return V, E, q, meshgen.model_params(ids, lam, mu)


def test_muscle_known_answers(oracle):
    """the fibre term of the muscle model (src/psi_muscle_fiber_F.cpp:5-10: 0.5 a d^T F^T F d) at
    F = diag(2, 4, 3) with d = e_y, a = 3: psi = 24, P = a (F d) d^T = 12 at (1, 1), d2psi = a at the
    three entries (i + 3, i + 3); added to the Neo-Hookean part exactly"""
    F = np.diag([2.0, 4.0, 3.0])
    p = np.array([0.0, 0.0, 3.0, 0.0, 1.0, 0.0])
    assert oracle.psi(5, F, p) == 24.0
    P = oracle.dpsi(5, F, p)
    Pref = np.zeros((3, 3))
    Pref[1, 1] = 12.0
    assert np.array_equal(P, Pref)
    H = oracle.d2psi(5, F, p, 1)
    Href = np.zeros((9, 9))
    for i in range(3):
        Href[i + 3, i + 3] = 3.0
    assert np.array_equal(H, Href)
    pn = np.array([3e5, 7e5, 3.0, 0.0, 1.0, 0.0])
    assert abs(oracle.psi(5, F, pn) - oracle.psi(0, F, pn[:2]) - 24.0) < 1e-9
    assert np.abs(oracle.dpsi(5, F, pn) - oracle.dpsi(0, F, pn[:2]) - Pref).max() < 1e-9
    assert np.abs(oracle.d2psi(5, F, pn, 1) - oracle.d2psi(0, F, pn[:2], 1) - Href).max() < 1e-9
    # a general direction: psi = a/2 |F d|^2 for d not along an axis
    d = np.array([0.3, -0.5, 0.8])
    pd = np.array([0.0, 0.0, 1.7, *d])
    assert abs(oracle.psi(5, F, pd) - 0.5 * 1.7 * np.sum((F @ d) ** 2)) < 1e-13


@pytest.mark.parametrize("model", MODELS + [5])
def test_mesh_fd(oracle, model):
    V, E, q, params = small_problem(model)
    dX, vol = oracle.dphidX_volume(V, E)
    nV = V.shape[0]
    g = oracle.tetmesh_dq(model, nV, E, q, dX, vol, params)
    H = oracle.tetmesh_dq2(model, nV, E, q, dX, vol, params, ref_layout=0)
    rng = np.random.default_rng(7)
    for _ in range(3):
        d = rng.standard_normal(q.size)
        h = 1e-6
        fp = oracle.tetmesh_q(model, nV, E, q + h * d, dX, vol, params)
        fm = oracle.tetmesh_q(model, nV, E, q - h * d, dX, vol, params)
        assert abs((fp - fm) / (2 * h) - g @ d) < 1e-5 * max(1.0, np.abs(g).sum() * np.abs(d).max())
        gp = oracle.tetmesh_dq(model, nV, E, q + h * d, dX, vol, params)
        gm = oracle.tetmesh_dq(model, nV, E, q - h * d, dX, vol, params)
        assert rel((gp - gm) / (2 * h), H @ d) < 1e-5
    # translation invariance: sum of the gradient over vertices vanishes, H annihilates translations
    assert np.abs(g.reshape(-1, 3).sum(0)).max() < 1e-9 * np.abs(g).max() * nV
    t = np.tile([1.0, 0.0, 0.0], nV)
    assert np.abs(H @ t).max() < 1e-9 * np.abs(H.data).max()


def test_fold_order_and_pattern(oracle):
    """values are left folds in ascending element order; the pattern is {(3vi+a, 3vj+b): vi, vj share a tet}"""
    V, E, q, params = small_problem(0, grid=(4, 3, 3))
    dX, vol = oracle.dphidX_volume(V, E)
    nV = V.shape[0]
    H = oracle.tetmesh_dq2(0, nV, E, q, dX, vol, params)
    dense = np.zeros((3 * nV, 3 * nV))
    pat = np.zeros((nV, nV), dtype=bool)
    for e in range(E.shape[0]):
        K = oracle.element_hessian(0, e, E, q, dX, vol, params)
        idx = (3 * E[e][:, None] + np.arange(3)).ravel()
        dense[np.ix_(idx, idx)] += K
        pat[np.ix_(E[e], E[e])] = True
    Hd = H.toarray()
    assert np.array_equal(Hd, dense)  # same fold order -> bit-identical
    bpat = np.kron(pat, np.ones((3, 3), dtype=bool))
    assert np.array_equal(H.toarray() != 0 | bpat, bpat)
    assert H.nnz == bpat.sum()


# ---- PSD ---------------------------------------------------------------------------------------
def test_psd_fix_spectral(oracle):
    rng = np.random.default_rng(3)
    for tol in (1e-6, 1e-3, 0.5):
        A = rng.standard_normal((12, 12))
        A = A + A.T
        w, U = np.linalg.eigh(A)
        ref = U @ np.diag(np.maximum(w, tol)) @ U.T
        assert rel(oracle.psd_fix(A, tol), ref) < 1e-12
    # fixed point on matrices whose eigenvalues are >= tol
    B = rng.standard_normal((12, 12))
    B = B @ B.T + np.eye(12)
    assert rel(oracle.psd_fix(B, 1e-3), B) < 1e-13


def test_element_psd_translations(oracle):
    V, E, q, params = small_problem(0, amp=0.3)
    dX, vol = oracle.dphidX_volume(V, E)
    tol = 1e-3
    for e in range(0, E.shape[0], 7):
        K = oracle.element_hessian(0, e, E, q, dX, vol, params, psd=True, psd_tol=tol)
        w = np.linalg.eigvalsh(K)
        assert w.min() > tol * (1 - 1e-6) - 1e-9 * np.abs(w).max()
        t = np.tile(np.eye(3), (4, 1)) / 2.0  # orthonormal translations
        assert rel(t.T @ K @ t, tol * np.eye(3)) < 1e-6 * np.abs(w).max() / tol


# ---- line search / Newton / direct solve -------------------------------------------------------
def test_linesearch_closed_form(oracle):
    """f(x) = x^2 at x = 1, d = -4 (g = 2): f(-3)=9 and f(-1)=1 fail the reference test
    f(x+a d) > f(x) + c g d = 1 - 8e-8; f(0) = 0 passes, so alpha = 0.25."""
    x, a, st = oracle.linesearch([1.0], [-4.0], [2.0], lambda x: float(x[0] ** 2))
    assert a == 0.25 and st == 0 and x[0] == 0.0
    # a NaN trial energy makes the comparison false: the reference accepts the step
    x, a, st = oracle.linesearch([1.0], [-4.0], [2.0], lambda x: float("nan") if x[0] < 0 else float(x[0] ** 2))
    assert a == 1.0 and x[0] == -3.0
    # cap: the step is applied even when the search gives up, status MAX_ITERATIONS_REACHED
    x, a, st = oracle.linesearch([0.0], [1.0], [-1.0], lambda x: 1.0 if x[0] != 0 else 0.0, max_it=3)
    assert st == 1 and a == 0.125 and x[0] == 0.125


def test_band_solve(oracle):
    V, E, q, params = small_problem(0)
    dX, vol = oracle.dphidX_volume(V, E)
    H = oracle.tetmesh_dq2(0, V.shape[0], E, q, dX, vol, params, psd=True)
    H = H + sp.eye(H.shape[0])
    b = np.random.default_rng(1).standard_normal(H.shape[0])
    x = oracle.band_solve(H, b)
    # normwise backward error at the level of FP64 rounding
    berr = np.linalg.norm(H @ x - b) / (sp.linalg.norm(H) * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr < 1e-14, berr
    assert rel(x, np.linalg.solve(H.toarray(), b)) < 1e-9


def test_newton_small_beam_converges(oracle):
    V, E = meshgen.beam_grid(6, 3, 3, h=0.05)
    dX, vol = oracle.dphidX_volume(V, E)
    lam, mu = meshgen.emu_to_lame(1e5, 0.4)
    params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
    fixed = np.zeros(V.size, np.uint8)
    fixed[(3 * meshgen.fixed_vertices(V)[:, None] + np.arange(3)).ravel()] = 1
    f = oracle.const_ext_force(V.shape[0], E, vol, 1000.0 * np.array([0, 0, -9.8]))
    # |g| < 1e-5 (absolute, initial |g| ~ 3.8); much tighter and the energy-based Armijo test stalls
    # on rounding of the total energy, as the reference's would
    q, it, conv, trace = oracle.newton_static(0, V.shape[0], E, dX, vol, params, V.ravel(), fixed, f, tol=1e-5,
                                              max_it=30)
    assert conv
    assert np.all(np.diff(trace[:, 0]) <= 1e-12 * np.abs(trace[0, 0]) + 1e-14)  # monotone objective
    assert trace[-1, 1] < 1e-5
    assert np.array_equal(q.reshape(-1, 3)[meshgen.fixed_vertices(V)], V[meshgen.fixed_vertices(V)])
    # quadratic convergence at the end
    gn = trace[:, 1]
    assert gn[-1] < 1e-3 * gn[-3]


def test_backward_euler_free_fall_translation(oracle):
    """Pins the oracle's backward-Euler inertia term (scale 1/dt^2, sign, mass): an unconstrained body at
    rest under the uniform load f = M g moves as a rigid translation by exactly dt^2 g in one step
    (elastic energy and gradient vanish on translations), and a second step continues with velocity
    dt g (x2 - x1 = 2 dt^2 g)"""
    V, E = meshgen.beam_grid(5, 3, 3, h=0.05)
    nV = V.shape[0]
    dX, vol = oracle.dphidX_volume(V, E)
    lam, mu = meshgen.emu_to_lame(1e6, 0.4)
    params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
    rho, dt = 1000.0, 0.01
    g = np.array([0.3, -1.1, -9.8])
    M = oracle.mass_matrix(nV, E, np.full(E.shape[0], rho), vol)
    f = M @ np.tile(g, nV)
    x0 = V.ravel().copy()
    x1, it, conv, _ = oracle.newton_be(0, nV, E, dX, vol, params, x0, x0, dt, rho, None, f, tol=1e-12, max_it=5)
    assert np.abs(x1 - x0 - dt * dt * np.tile(g, nV)).max() < 1e-12
    v1 = (x1 - x0) / dt
    qp = x1 + dt * v1
    x2, _, _, _ = oracle.newton_be(0, nV, E, dX, vol, params, qp, qp, dt, rho, None, f, tol=1e-12, max_it=5)
    assert np.abs(x2 - x1 - 2 * dt * dt * np.tile(g, nV)).max() < 1e-12


def test_psd_fix_ql_matches_eigh(oracle):
    """The CPU baseline's tridiagonalisation + implicit QR simple_psd_fix against numpy's eigh spectral
    function and the checker's Jacobi version: random symmetric matrices (indefinite, wide scale
    ranges, repeated eigenvalues, already-diagonal and already-tridiagonal inputs) and real element
    Hessians of strongly distorted tets."""
    rng = np.random.default_rng(3)

    def spectral(A, tol):
        w, V = np.linalg.eigh(A)
        return (V * np.maximum(w, tol)) @ V.T

    cases = []
    for k in range(20):
        A = rng.standard_normal((12, 12))
        cases.append(((A + A.T) * 10.0 ** rng.uniform(-4, 6), 10.0 ** rng.uniform(-3, 2)))
    Q, _ = np.linalg.qr(rng.standard_normal((12, 12)))
    w = np.array([3.0, 3.0, 3.0, -1.0, -1.0, 0.0, 0.0, 0.0, 5.0, 7.0, 7.0, 1e-9])
    cases.append(((Q * w) @ Q.T, 1e-3))  # repeated eigenvalues
    cases.append((np.diag(rng.standard_normal(12)), 0.1))  # already diagonal
    off = rng.standard_normal(11)
    T = np.diag(rng.standard_normal(12)) + np.diag(off, 1) + np.diag(off, -1)
    T[4, 5] = T[5, 4] = 0.0  # a split tridiagonal
    cases.append((T, 0.2))
    V, E = meshgen.beam_grid(4, 3, 3, h=0.02)
    q = (V + 0.3 * 0.02 * rng.standard_normal(V.shape)).ravel()
    dX, vol = oracle.dphidX_volume(V, E)
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
    for e in range(0, E.shape[0], 7):
        cases.append((oracle.element_hessian(0, e, E, q, dX, vol, params), 1e-3))
    for A, tol in cases:
        ref = spectral(A, tol)
        got = oracle.psd_fix_ql(A, tol)
        scale = max(np.abs(A).max(), tol)
        assert np.abs(got - ref).max() < 1e-13 * scale * 12
        assert np.abs(got - oracle.psd_fix(A, tol)).max() < 1e-12 * scale
        assert np.array_equal(got, got.T)


def test_timed_dq2_matches_checker(oracle):
    """the CPU baseline's timing mode (QL simple_psd_fix, triplets at 144 e, OpenMP over elements)
    assembles the checker's matrix: same pattern, values to rounding, identical bits for 1 and 4
    threads (the triplet list is in element order whatever the thread count)"""
    import scipy.sparse as sp

    V, E = meshgen.beam_grid(8, 4, 4, h=0.02)
    rng = np.random.default_rng(5)
    q = (V + 0.2 * 0.02 * rng.standard_normal(V.shape)).ravel()
    dX, vol = oracle.dphidX_volume(V, E)
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
    nV = V.shape[0]
    H_ref = oracle.tetmesh_dq2(0, nV, E, q, dX, vol, params, psd=True, psd_tol=1e-3)
    H1, t1 = oracle.tetmesh_dq2_timed(0, nV, E, q, dX, vol, params, psd=True, psd_tol=1e-3, threads=1, fast=False)
    H4, t4 = oracle.tetmesh_dq2_timed(0, nV, E, q, dX, vol, params, psd=True, psd_tol=1e-3, threads=4, fast=False)
    assert np.array_equal(H1.indptr, H_ref.indptr) and np.array_equal(H1.indices, H_ref.indices)
    assert sp.linalg.norm(H1 - H_ref) / sp.linalg.norm(H_ref) < 1e-13
    assert np.array_equal(H1.data.view(np.uint64), H4.data.view(np.uint64))
    assert t1[0] > 0 and t1[1] > 0

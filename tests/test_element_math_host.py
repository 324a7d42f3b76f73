"""The device element math (bartels_amd/csrc/elements.h: closed-form blocks, Piola stress,
energy; FP64 and the FP32 precision form) compiled for the host and checked tet by tet against
the oracle's dense reference algorithm (B^T d2psi B with the generated-style 9x9).  Runs on CPU."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from bartels_amd import meshgen
from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def eh(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("eh") / "eh.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
                    "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "element_harness.cpp"), "-o", out],
                   check=True)
    L = ctypes.CDLL(out)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.eh_blocks.argtypes = [i, i, vp, vp, d, d, d, i, vp]
    L.eh_grad.argtypes = [i, i, vp, vp, d, d, d, vp]
    L.eh_energy.argtypes = [i, i, vp, vp, d, d, d]
    L.eh_energy.restype = d
    L.eh_psd_blocks.argtypes = [i, i, vp, vp, d, d, d, i, d, vp]
    L.eh_psd_blocks.restype = i
    L.eh_eig9.argtypes = [i, vp, vp, vp]
    L.eh_psd_low9.argtypes = [i, vp, d, vp]
    L.eh_psd_low9.restype = i
    L.eh_grad_g.argtypes = [i, vp, vp, vp, d, d, d, vp]
    return L


def tets(model, n=24, seed=0):
    """single tets cut from a jittered beam, deformed by a random stretch (distinct singular values)"""
    V, E = meshgen.beam_grid(5, 4, 4, h=0.02)
    V = meshgen.jitter(V, E, 0.02, 0.2, seed=seed + 1)
    rng = np.random.default_rng(seed)
    out = []
    for e in rng.choice(E.shape[0], n, replace=False):
        X = V[E[e]]
        Q1, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        Q2, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        A = Q1 @ np.diag([1.25, 1.0, 0.8]) @ Q2
        if np.linalg.det(A) < 0:
            A = -A
        x = X @ A.T + 3.0 + 0.002 * rng.standard_normal(X.shape)  # offset: metres-scale coordinates
        lam, mu = meshgen.emu_to_lame(rng.uniform(1e5, 1e7), rng.uniform(0.3, 0.48))
        p = meshgen.model_params(np.array([model]), lam, mu)[0]
        out.append((X, x, p))
    return out


def single_tet(oracle, X):
    E = np.array([[0, 1, 2, 3]], dtype=np.int32)
    dX, vol = oracle.dphidX_volume(X, E)
    return E, dX, vol


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("ref", [1, 0])
def test_blocks_match_dense_reference(eh, oracle, model, ref):
    if ref == 0 and model not in (2, 3):
        pytest.skip("layout only affects stretch models")
    for X, x, p in tets(model):
        E, dX, vol = single_tet(oracle, X)
        q = x.ravel()
        K_or = oracle.element_hessian(model, 0, E, q, dX, vol, p[None, :], ref_layout=ref)
        phi = np.ascontiguousarray(dX[0])
        xx = np.ascontiguousarray(q)
        for fp32, tol in ((0, 1e-12), (1, 2e-5)):
            K = np.zeros(144)
            eh.eh_blocks(model, fp32, phi.ctypes.data, xx.ctypes.data, vol[0], p[0], p[1], ref, K.ctypes.data)
            K = K.reshape(12, 12, order="F")
            err = np.abs(K - K_or).max() / np.abs(K_or).max()
            assert err < tol, (model, ref, fp32, err)


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
def test_gradient_and_energy_match_reference(eh, oracle, model):
    for X, x, p in tets(model, seed=5):
        E, dX, vol = single_tet(oracle, X)
        q = x.ravel()
        g_or = oracle.tetmesh_dq(model, 4, E, q, dX, vol, p[None, :])
        e_or = oracle.tetmesh_q(model, 4, E, q, dX, vol, p[None, :])
        phi = np.ascontiguousarray(dX[0])
        for fp32, tg, te in ((0, 1e-12, 1e-12), (1, 2e-5, 5e-6)):
            g = np.zeros(12)
            eh.eh_grad(model, fp32, phi.ctypes.data, q.ctypes.data, vol[0], p[0], p[1], g.ctypes.data)
            e = eh.eh_energy(model, fp32, phi.ctypes.data, q.ctypes.data, vol[0], p[0], p[1])
            assert np.abs(g - g_or).max() / np.abs(g_or).max() < tg, (model, fp32)
            assert abs(e - e_or) / abs(e_or) < te, (model, fp32, e, e_or)


def test_fp32_energy_keeps_relative_accuracy_near_rest(eh, oracle):
    """the precision form's energy is formed from G = F - I and J - 1 in FP64: at strain 1e-4 it still
    matches the FP64 energy to ~1e-6 relative (a plain FP32 tr(F^T F) - 3 would lose everything)"""
    rng = np.random.default_rng(9)
    V, E = meshgen.beam_grid(3, 3, 3, h=0.025)
    X = V[E[0]]
    Ex = np.array([[0, 1, 2, 3]], dtype=np.int32)
    dX, vol = oracle.dphidX_volume(X, Ex)
    x = X + 2.0 + 1e-4 * 0.025 * rng.standard_normal(X.shape)
    p = np.array([3.4e5, 3.1e6])
    e_or = oracle.tetmesh_q(0, 4, Ex, x.ravel(), dX, vol, p[None, :])
    q = np.ascontiguousarray(x.ravel())
    phi = np.ascontiguousarray(dX[0])
    e32 = eh.eh_energy(0, 1, phi.ctypes.data, q.ctypes.data, vol[0], p[0], p[1])
    assert abs(e32 - e_or) / abs(e_or) < 5e-3


@pytest.mark.parametrize("model", [0, 1, 2, 4])
def test_psd_decomposition_matches_12x12_eigensolve(eh, oracle, model):
    """simple_psd_fix through the reduced 9x9 (QR of Phi, Cholesky test, the clamped record M' by
    inverse iteration for k <= 3 eigenvalues below tol and by the full tridiagonal-QL eigensolver for
    k > 3, blocks from M' and Q) equals the oracle's 12x12 Jacobi eigen-clamp, both on tets that need
    clamping and on ones that do not, and is exactly symmetric; the large tolerance forces k > 3"""
    rng = np.random.default_rng(11 + model)
    V, E = meshgen.beam_grid(5, 4, 4, h=0.02)
    formats = set()
    for e in rng.choice(E.shape[0], 30, replace=False):
        X = V[E[e]]
        x = X + 0.3 * 0.02 * rng.standard_normal(X.shape)  # strong distortion: indefinite element Hessians
        lam, mu = meshgen.emu_to_lame(2e6, 0.45)
        p = meshgen.model_params(np.array([model]), lam, mu)[0]
        Ex = np.array([[0, 1, 2, 3]], dtype=np.int32)
        dX, vol = oracle.dphidX_volume(X, Ex)
        if vol[0] <= 0:
            continue
        q = np.ascontiguousarray(x.ravel())
        phi = np.ascontiguousarray(dX[0])
        for tol in (1e-3, 10.0, 3e4):
            K_or = oracle.element_hessian(model, 0, Ex, q, dX, vol, p[None, :], psd=True, psd_tol=tol)
            for fp32, bound in ((0, 1e-12), (1, 1e-4)):
                K = np.zeros(144)
                fmt = eh.eh_psd_blocks(model, fp32, phi.ctypes.data, q.ctypes.data, vol[0], p[0], p[1], 1, tol,
                                       K.ctypes.data)
                formats.add(fmt)
                K = K.reshape(12, 12, order="F")
                assert np.array_equal(K, K.T)  # blocks (m, a) are the exact transposes of (a, m)
                err = np.abs(K - K_or).max() / np.abs(K_or).max()
                assert err < bound, (model, tol, fp32, fmt, err)
    assert 4 in formats and formats & {1, 2, 3}, formats


def _eig9(eh, A, fp32=0):
    d = np.zeros(9)
    Z = np.zeros(81)
    A = np.asfortranarray(A)
    eh.eh_eig9(fp32, A.ctypes.data, d.ctypes.data, Z.ctypes.data)
    return d, Z.reshape(9, 9, order="F")


def test_sym9_eigen_against_closed_forms_and_eigh(eh):
    """the tridiagonal-QL eigensolver: diagonal and block-diagonal inputs (known spectra), repeated
    and clustered eigenvalues, rank-deficient and indefinite matrices, wide scale ranges; checked by
    orthonormality, residual |A Z - Z diag(d)| and the sorted spectrum of numpy's eigh"""
    rng = np.random.default_rng(5)
    cases = [np.diag(np.arange(9.0) - 4.0), np.zeros((9, 9)), np.eye(9) * 3.0]
    B = np.zeros((9, 9))
    B[:3, :3] = [[2, 1, 0], [1, 2, 0], [0, 0, 5]]  # eigenvalues 1, 3, 5
    B[3:, 3:] = np.diag([-1.0, -1.0, 7.0, 7.0, 7.0, 0.0])
    cases.append(B)
    for trial in range(200):
        Q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
        kind = trial % 5
        if kind == 0:
            lam = rng.standard_normal(9) * 10 ** rng.uniform(-3, 4)
        elif kind == 1:
            lam = np.repeat(rng.standard_normal(3), 3)  # triple eigenvalues
        elif kind == 2:
            lam = 1.0 + 1e-9 * rng.standard_normal(9)  # tight cluster
        elif kind == 3:
            lam = np.concatenate([rng.standard_normal(4) * 1e3, np.zeros(5)])  # rank 4
        else:
            lam = np.sign(rng.standard_normal(9)) * 10 ** rng.uniform(-6, 6, 9)  # wide dynamic range
        cases.append((Q * lam) @ Q.T)
    for A in cases:
        A = 0.5 * (A + A.T)
        nA = max(np.abs(A).max(), 1e-300)
        d, Z = _eig9(eh, A)
        assert np.abs(Z.T @ Z - np.eye(9)).max() < 5e-15 * 9
        assert np.abs(A @ Z - Z * d).max() <= 5e-15 * 9 * nA + 1e-300
        assert np.abs(np.sort(d) - np.linalg.eigvalsh(A)).max() <= 5e-15 * 9 * nA + 1e-300
        d32, Z32 = _eig9(eh, A, fp32=1)
        assert np.abs(np.sort(d32) - np.linalg.eigvalsh(A)).max() <= 3e-6 * nA + 1e-30
        assert np.abs(Z32.T @ Z32 - np.eye(9)).max() < 3e-6


def test_psd_low_rank_by_inverse_iteration(eh):
    """the clamp correction for k <= 3 eigenvalues below tol without the eigenvector matrix (QL
    eigenvalues, inverse iteration on the tridiagonal form, reflectors applied back): W W^T equals
    sum (tol - lambda) v v^T from numpy's eigh, on isolated, clustered (1e-9 and exactly repeated),
    zero and wide-range low eigenvalues and on diagonal and block-diagonal inputs (split tridiagonal
    forms); for k > 3 it declines"""
    rng = np.random.default_rng(17)
    cases = []
    for trial in range(300):
        Q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
        hi = rng.uniform(1.0, 1e3, 9)
        kind = trial % 6
        k = 1 + trial % 3
        lam = hi.copy()
        if kind == 0:
            lam[:k] = rng.uniform(-1.0, 0.9, k)  # isolated
        elif kind == 1:
            lam[:k] = 0.3 + 1e-9 * rng.standard_normal(k)  # tight cluster
        elif kind == 2:
            lam[:k] = 0.0  # exactly repeated (rigid-rotation-like null modes)
        elif kind == 3:
            lam[:k] = -10 ** rng.uniform(-8, 3, k)  # wide range
        elif kind == 4:
            lam[:k] = rng.uniform(-1.0, 0.9, k)
            lam[k] = 1.0 + 1e-9  # a partner just above tol
        else:
            lam[:k] = rng.uniform(-1.0, 0.9, k)
            Q = np.eye(9)[rng.permutation(9)]  # diagonal input: every off-diagonal of T is zero
        cases.append(((Q * lam) @ Q.T, 1.0))
    B = np.zeros((9, 9))
    B[:3, :3] = [[2, 1, 0], [1, 2, 0], [0, 0, 5]]
    B[3:, 3:] = np.diag([-1.0, 9.0, 7.0, 7.0, 7.0, 0.5])
    cases.append((B, 1.5))  # three below tol across decoupled blocks (eigenvalues 1, -1, 0.5)
    for A, tol in cases:
        A = 0.5 * (A + A.T)
        w, V = np.linalg.eigh(A)
        sel = w < tol
        dM = (V[:, sel] * (tol - w[sel])) @ V[:, sel].T
        nA = np.abs(A).max()
        for fp32, bound in ((0, 2e-14), (1, 2e-5)):
            W = np.zeros(27)
            k = eh.eh_psd_low9(fp32, np.asfortranarray(A).ctypes.data, tol, W.ctypes.data)
            if fp32:  # an eigenvalue within FP32 rounding of tol may count either way
                if k > 3:
                    continue
            else:
                assert k == sel.sum()
            W = W.reshape(3, 9)[:k].T
            err = np.abs(W @ W.T - dM).max() / nA
            assert err < bound, (fp32, k, err)
    A = np.diag(np.r_[np.zeros(4), np.ones(5) * 5.0])
    assert eh.eh_psd_low9(0, np.asfortranarray(A).ctypes.data, 1.0, np.zeros(27).ctypes.data) == 4


@pytest.mark.parametrize("model", [0, 1])
def test_fp32_stress_from_displacements_small_strain(eh, oracle, model):
    """the precision form's small-strain path: G = F - I from displacement edges (FP64), stress from G
    in FP32.  At strain 1e-5 its gradient stays within 1e-5 of the FP64 oracle, where the path through
    F rounded to FP32 loses ~1e-2"""
    rng = np.random.default_rng(21 + model)
    worst_g, worst_f = 0.0, 0.0
    for _ in range(20):
        base = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])  # positive signed volume
        X = 3.0 + 0.02 * (base + 0.1 * rng.standard_normal((4, 3)))
        x = X + 1e-5 * 0.02 * rng.standard_normal((4, 3))
        lam, mu = meshgen.emu_to_lame(2e6, 0.45)
        p = meshgen.model_params(np.array([model]), lam, mu)[0]
        Ex = np.array([[0, 1, 2, 3]], dtype=np.int32)
        dX, vol = oracle.dphidX_volume(X, Ex)
        assert vol[0] > 0
        q = np.ascontiguousarray(x.ravel())
        g_or = oracle.tetmesh_dq(model, 4, Ex, q, dX, vol, p[None, :])
        phi = np.ascontiguousarray(dX[0])
        Xr = np.ascontiguousarray(X.ravel())
        g = np.zeros(12)
        eh.eh_grad_g(model, phi.ctypes.data, q.ctypes.data, Xr.ctypes.data, vol[0], p[0], p[1], g.ctypes.data)
        gf = np.zeros(12)
        eh.eh_grad(model, 1, phi.ctypes.data, q.ctypes.data, vol[0], p[0], p[1], gf.ctypes.data)
        n = np.linalg.norm(g_or)
        worst_g = max(worst_g, np.linalg.norm(g - g_or) / n)
        worst_f = max(worst_f, np.linalg.norm(gf - g_or) / n)
    print(f"model {model}: FP32 gradient error at strain 1e-5, from G {worst_g:.1e}, from F {worst_f:.1e}")
    assert worst_g < 1e-5

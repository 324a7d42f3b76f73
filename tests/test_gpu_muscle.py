"""Neo-Hookean plus a muscle fibre (BH_NEOHOOKEAN_MUSCLE: the generic_elastic functor of
matlab/linear_tetmesh_neohookean_muscle_{q,dq,dq2}.cpp) against the oracle: energy, gradient, Hessian
with and without simple_psd_fix, through the reference-shaped mirror functions; the FP32 element path;
and a static Newton solve of a fibre-loaded beam against the oracle's direct-solve Newton."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


def _problem(seed=0):
    V, E = meshgen.beam_grid(10, 5, 5, h=0.02)
    V = meshgen.jitter(V, E, 0.02, 0.15, seed=seed + 2)
    nT = E.shape[0]
    rng = np.random.default_rng(seed)
    lam, mu = meshgen.emu_to_lame(np.exp(np.log(1e5) + np.log(100.0) * rng.random(nT)), 0.3 + 0.15 * rng.random(nT))
    params_nh = meshgen.model_params(np.zeros(nT, np.uint8), lam, mu)
    d = rng.standard_normal((nT, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    params_a = (lam * (0.2 + rng.random(nT)))[:, None]
    q = (V + 0.05 * 0.02 * rng.standard_normal(V.shape)).ravel()
    return V, E, q, d, params_nh, params_a


def test_muscle_parity(oracle):
    from bartels_amd import tetmesh as tm

    V, E, q, d, params_nh, params_a = _problem()
    nV = V.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    P6 = tm.muscle_params(d, params_nh, params_a)
    e_gpu = tm.linear_tetmesh_neohookean_muscle_q(V, E, q, dX, vol, d, params_nh, params_a)
    g_gpu = tm.linear_tetmesh_neohookean_muscle_dq(V, E, q, dX, vol, d, params_nh, params_a)
    e_or = oracle.tetmesh_q(5, nV, E, q, dX, vol, P6)
    g_or = oracle.tetmesh_dq(5, nV, E, q, dX, vol, P6)
    # the fibre term is present: the muscle energy exceeds the Neo-Hookean one
    assert e_or > oracle.tetmesh_q(0, nV, E, q, dX, vol, params_nh) * (1 + 1e-3)
    print(f"muscle energy rel {abs(e_gpu - e_or) / abs(e_or):.2e}, gradient rel "
          f"{np.linalg.norm(g_gpu - g_or) / np.linalg.norm(g_or):.2e}")
    assert abs(e_gpu - e_or) < 1e-12 * abs(e_or)
    assert np.linalg.norm(g_gpu - g_or) < 1e-12 * np.linalg.norm(g_or)
    for cb, kw in ((None, {}), (tm.simple_psd_fix(1e-3), dict(psd=True, psd_tol=1e-3)),
                   (tm.simple_psd_fix(1e2), dict(psd=True, psd_tol=1e2))):
        H_gpu = tm.linear_tetmesh_neohookean_muscle_dq2(V, E, q, dX, vol, d, params_nh, params_a, callback=cb)
        H_or = oracle.tetmesh_dq2(5, nV, E, q, dX, vol, P6, **kw)
        eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
        print(f"  Hessian {kw or 'no PSD'}: rel {eH:.2e}")
        assert eH < 1e-12


def test_muscle_precision_form(oracle):
    """the FP32 element path (params nT x 6 in FP32) within the precision form's tolerances"""
    import torch

    from bartels_amd import tetmesh as tm

    V, E, q, d, params_nh, params_a = _problem(1)
    nV = V.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    P6 = tm.muscle_params(d, params_nh, params_a)
    mesh = tm.TetMesh(nV, E, dX, vol, V=V)
    try:
        dq = torch.tensor(q, device="cuda")
        p32 = torch.tensor(np.asfortranarray(P6).ravel(order="F"), dtype=torch.float32, device="cuda")
        e = torch.empty(1, dtype=torch.float64, device="cuda")
        g = torch.empty(3 * nV, dtype=torch.float64, device="cuda")
        H = torch.empty(mesh.nnz, dtype=torch.float32, device="cuda")
        mesh.q_f32_device(tm.NEOHOOKEAN_MUSCLE, dq, p32, e)
        mesh.dq_f32_device(tm.NEOHOOKEAN_MUSCLE, dq, p32, g)
        mesh.dq2_f32_device(tm.NEOHOOKEAN_MUSCLE, dq, p32, H, callback=tm.simple_psd_fix(1e-3))
        torch.cuda.synchronize()
        e_or = oracle.tetmesh_q(5, nV, E, q, dX, vol, P6)
        g_or = oracle.tetmesh_dq(5, nV, E, q, dX, vol, P6)
        H_or = oracle.tetmesh_dq2(5, nV, E, q, dX, vol, P6, psd=True, psd_tol=1e-3)
        rowptr, colind = mesh.pattern()
        Hg = sp.csr_matrix((H.cpu().numpy().astype(np.float64), colind, rowptr), shape=H_or.shape)
        ee = abs(e.item() - e_or) / abs(e_or)
        eg = np.linalg.norm(g.cpu().numpy() - g_or) / np.linalg.norm(g_or)
        eH = sp.linalg.norm(Hg - H_or) / sp.linalg.norm(H_or)
        print(f"FP32 muscle: energy {ee:.2e}, gradient {eg:.2e}, Hessian (PSD) {eH:.2e}")
        assert ee < 1e-6 and eg < 1e-5 and eH < 1e-5
    finally:
        mesh.close()


def test_muscle_newton_matches_oracle(oracle):
    """a clamped beam whose fibres pull it out of its rest shape (the fibre energy is not zero at rest):
    the device Newton (FP64, block-Jacobi PCG) reaches the oracle's direct-solve equilibrium"""
    import torch

    from bartels_amd import _lib
    from bartels_amd import tetmesh as tm

    V, E = meshgen.beam_grid(12, 4, 4, h=0.025)
    nV, nT = V.shape[0], E.shape[0]
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    nh = meshgen.model_params(np.zeros(nT, np.uint8), lam, mu)
    d = np.tile([1.0, 0.0, 0.0], (nT, 1))
    P6 = tm.muscle_params(d, nh, np.full(nT, 0.2 * lam))
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    mesh = tm.TetMesh(nV, E, dX, vol)
    fixed = np.zeros(3 * nV, np.uint8)
    fixed[(3 * meshgen.fixed_vertices(V)[:, None] + np.arange(3)).ravel()] = 1
    q0 = V.ravel().copy()
    q_or, _, conv, _ = oracle.newton_static(5, nV, E, dX, vol, P6, q0, fixed, None, psd=True, tol=1e-8, max_it=30)
    assert conv
    lib = _lib.load()
    o = _lib.BhNewtonOpts()
    lib.bh_newton_default_opts(ctypes.byref(o))
    o.model = tm.NEOHOOKEAN_MUSCLE
    o.psd = _lib.BhPsd(1, 1e-3)
    o.gradient_tol = 1e-8
    o.max_iterations = 30
    o.pcg_tol = 1e-12
    dq = torch.tensor(q0, device="cuda")
    dpar = torch.tensor(np.asfortranarray(P6).ravel(order="F"), device="cuda")
    dfixed = torch.tensor(fixed, device="cuda")
    rep = _lib.BhNewtonReport()
    try:
        st = lib.bh_newton_static_device(mesh.handle, ctypes.byref(o), ctypes.c_void_p(dpar.data_ptr()),
                                         ctypes.c_void_p(dfixed.data_ptr()), None, ctypes.c_void_p(dq.data_ptr()),
                                         ctypes.byref(rep), None,
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == _lib.BH_OK and rep.converged
        u = np.linalg.norm(q_or - q0)
        err = np.linalg.norm(dq.cpu().numpy() - q_or) / u
        print(f"muscle Newton: displacement {u:.3e}, device vs oracle {err:.2e} ({rep.iterations} iterations)")
        assert u > 1e-4 and err < 1e-8
    finally:
        mesh.close()

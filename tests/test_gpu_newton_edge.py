"""Edge cases of the device Newton step against the oracle: the reference line search's quirks
(src/linesearch_backtracking_bisection.cpp:21-29: a NaN trial energy is accepted, the step is applied
when the search hits its cap), and the Jacobi fallback of the factor solvers on breakdown (a singular
or non-finite factor makes r.u vanish or turn NaN)."""
import ctypes

import numpy as np
import pytest

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu

H = 0.025


def _newton(mesh, model, params, q0, fixed, f, max_it=0, gtol=0.0, relative=0, pcg_tol=1e-13, precond=0,
            handle=None, precision=0, fallback=1, max_ls=100, reject_nonfinite=0, pcg_max=20000):
    import torch

    from bartels_amd import _lib

    lib = _lib.load()
    o = _lib.BhNewtonOpts()
    lib.bh_newton_default_opts(ctypes.byref(o))
    o.model = model
    o.gradient_tol = gtol
    o.relative_tol = relative
    o.max_iterations = max_it
    o.max_iterations_ls = max_ls
    o.reject_nonfinite = reject_nonfinite
    o.pcg_tol = pcg_tol
    o.pcg_max_iterations = pcg_max
    o.precond = precond
    o.precond_handle = handle
    o.element_precision = precision
    o.jacobi_fallback = fallback
    dq = torch.tensor(q0, device="cuda")
    dparams = torch.tensor(np.asfortranarray(params).ravel(order="F"), device="cuda")
    dfixed = torch.tensor(fixed, device="cuda")
    df = torch.tensor(f, device="cuda")
    rep = _lib.BhNewtonReport()
    trace = np.zeros(4 * (max_it + 1))
    st = lib.bh_newton_static_device(mesh.handle, ctypes.byref(o), ctypes.c_void_p(dparams.data_ptr()),
                                     ctypes.c_void_p(dfixed.data_ptr()), ctypes.c_void_p(df.data_ptr()),
                                     ctypes.c_void_p(dq.data_ptr()), ctypes.byref(rep),
                                     trace.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, dq.cpu().numpy(), rep, trace.reshape(-1, 4)


def _small_beam(oracle, model, direction, scale_lengths):
    """a 4x2x2-vertex beam clamped at x = 0, loaded so that the first Newton step from rest moves the
    free end by scale_lengths beam lengths (linearised): f = s * (gravity-like load along direction)"""
    import scipy.sparse.linalg as spla

    from bartels_amd import tetmesh as tm

    V, E = meshgen.beam_grid(4, 2, 2, h=H)
    nV, nT = V.shape[0], E.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    fixed = np.zeros(3 * nV, np.uint8)
    fixed[(3 * meshgen.fixed_vertices(V)[:, None] + np.arange(3)).ravel()] = 1
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.full(nT, model, np.uint8), lam, mu)
    q0 = V.ravel().copy()
    g = oracle.const_ext_force(nV, E, vol, 9.8e3 * np.asarray(direction, dtype=np.float64))
    K = oracle.tetmesh_dq2(model, nV, E, q0, dX, vol, params).tocsc()
    free = fixed == 0
    d = np.zeros(3 * nV)
    d[free] = spla.spsolve(K[free][:, free], g[free])
    f = g * (scale_lengths * 3 * H / np.abs(d).max())
    mesh = tm.TetMesh(nV, E, dX, vol)
    return mesh, V, E, dX, vol, params, q0, fixed, f


def test_linesearch_accepts_nan_trial(oracle):
    """Neo-Hookean beam pushed along -x so that the full Newton step inverts tets (log J of a negative J:
    NaN trial energies at alpha = 1 and 0.5).  Reference semantics: NaN > f + c g.d is false, so the
    step alpha = 1 is taken and the objective becomes NaN -- the device and the oracle agree on alpha and
    on q.  With reject_nonfinite = 1 the device backtracks to the first finite trial that passes the
    Armijo test instead (alpha = 0.25 here, where the oracle's energies are finite and decreasing)."""
    mesh, V, E, dX, vol, params, q0, fixed, f = _small_beam(oracle, 0, (-1.0, 0.0, 0.0), 2.0)
    nV = V.shape[0]
    try:
        q_or, it_or, _, tr_or = oracle.newton_static(0, nV, E, dX, vol, params, q0, fixed, f, tol=0.0, max_it=0)
        assert it_or == 1 and tr_or[0, 2] == 1.0
        assert np.isnan(oracle.tetmesh_q(0, nV, E, q_or, dX, vol, params))  # the accepted step inverted tets
        st, q, rep, tr = _newton(mesh, 0, params, q0, fixed, f)
        err = np.linalg.norm(q - q_or) / np.linalg.norm(q_or - q0)
        print(f"NaN trial: device alpha {tr[0, 2]}, oracle alpha {tr_or[0, 2]}, q difference {err:.2e}, "
              f"device energy after {mesh.q(0, q, params):.3e}")
        assert st == 4  # BH_MAXIT: tolerance 0, one iteration
        assert tr[0, 2] == 1.0 and rep.ls_iterations_total == 0
        assert err < 1e-10
        assert np.isnan(mesh.q(0, q, params))
        # the textbook variant rejects non-finite trials
        e_half = oracle.tetmesh_q(0, nV, E, q0 + 0.5 * (q_or - q0), dX, vol, params)
        e_quarter = oracle.tetmesh_q(0, nV, E, q0 + 0.25 * (q_or - q0), dX, vol, params)
        assert np.isnan(e_half) and np.isfinite(e_quarter)
        st, q2, rep2, tr2 = _newton(mesh, 0, params, q0, fixed, f, reject_nonfinite=1)
        assert tr2[0, 2] == 0.25 and rep2.ls_iterations_total == 2
        err2 = np.linalg.norm(q2 - (q0 + 0.25 * (q_or - q0))) / np.linalg.norm(q_or - q0)
        assert err2 < 1e-10
        assert np.isfinite(mesh.q(0, q2, params))
    finally:
        mesh.close()


@pytest.mark.parametrize("max_ls", [2, 3])
def test_linesearch_cap_applies_step(oracle, max_ls):
    """StVK beam bent by a load whose full Newton step overshoots badly (energies at alpha = 1, 0.5, 0.25
    all fail the Armijo test; 0.125 passes).  With max_iterations_ls = 2 the search stops at its cap and
    the reference still applies alpha = 0.25 (MAX_ITERATIONS_REACHED); with 3 the accepted alpha = 0.125
    is reached on the cap itself.  Device and oracle agree on alpha, the backtrack count and q."""
    mesh, V, E, dX, vol, params, q0, fixed, f = _small_beam(oracle, 1, (0.0, 0.0, -1.0), 5.0)
    nV = V.shape[0]
    try:
        q_or, it_or, _, tr_or = oracle.newton_static(1, nV, E, dX, vol, params, q0, fixed, f, tol=0.0, max_it=0,
                                                     max_ls=max_ls)
        alpha = 0.5 ** max_ls
        assert tr_or[0, 2] == alpha and tr_or[0, 3] == max_ls
        st, q, rep, tr = _newton(mesh, 1, params, q0, fixed, f, max_ls=max_ls)
        d = (q_or - q0) / alpha
        err = np.linalg.norm(q - q_or) / np.linalg.norm(alpha * d)
        print(f"cap {max_ls}: device alpha {tr[0, 2]}, backtracks {rep.ls_iterations_total}; q difference {err:.2e}")
        assert tr[0, 2] == alpha and rep.ls_iterations_total == max_ls
        assert err < 1e-10
        if max_ls == 2:  # the capped step does not decrease the objective: it is applied all the same
            e0 = oracle.tetmesh_q(1, nV, E, q0, dX, vol, params) - f @ q0
            e1 = oracle.tetmesh_q(1, nV, E, q_or, dX, vol, params) - f @ q_or
            assert e1 > e0
    finally:
        mesh.close()


@pytest.mark.parametrize("kind", ["zero", "nan"])
@pytest.mark.parametrize("precision", [0, 1])
def test_fallback_on_factor_breakdown(oracle, kind, precision):
    """A factor preconditioner whose blocks are all zero (M^-1 = 0, so r.u = 0 in the first CG
    iteration) or NaN: CG breaks down rather than hitting its iteration cap.  With the Jacobi
    fallback every Newton direction is redone with block Jacobi and the solve reaches the oracle's
    equilibrium (10x10x40 beam under gravity); without it the direction is zero and Newton makes no
    progress, reported as one failed inner solve per iteration."""
    import torch

    from bartels_amd import _lib
    from bartels_amd import tetmesh as tm

    p = meshgen.config1()
    nV = p.V.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
    mesh = tm.TetMesh(nV, p.E, dX, vol, V=p.V)
    fixed = np.zeros(3 * nV, np.uint8)
    fixed[(3 * p.fixed[:, None] + np.arange(3)).ravel()] = 1
    q0 = p.V.ravel().copy()
    f = oracle.const_ext_force(nV, p.E, vol, p.density * p.gravity)
    q_or, _, conv, _ = oracle.newton_static(0, nV, p.E, dX, vol, p.params, q0, fixed, f, tol=1e-7, max_it=20)
    assert conv
    lib = _lib.load()
    P = ctypes.c_void_p()
    assert lib.bh_precond_create(mesh.handle, ctypes.byref(P)) == 0
    try:
        ncol, nlow = ctypes.c_int(), ctypes.c_int64()
        assert lib.bh_precond_info(P, ctypes.byref(ncol), ctypes.byref(nlow)) == 0
        fill = 0.0 if kind == "zero" else float("nan")
        lower = torch.full((9 * nlow.value,), fill, dtype=torch.float32, device="cuda")
        dinv = torch.full((6 * nV,), fill, dtype=torch.float32, device="cuda")
        assert lib.bh_precond_set_factor_device(P, ctypes.c_void_p(lower.data_ptr()), ctypes.c_void_p(dinv.data_ptr()),
                                                None) == 0
        torch.cuda.synchronize()
        u = np.linalg.norm(q_or - q0)
        st, q, rep, _ = _newton(mesh, 0, p.params, q0, fixed, f, max_it=30, gtol=1e-7, relative=1, pcg_tol=1e-10,
                                precond=3, handle=P, precision=precision)
        err = np.linalg.norm(q - q_or) / u
        print(f"{kind} factor, precision {precision}: status {st}, newton {rep.iterations}, fallbacks "
              f"{rep.pcg_fallbacks}, pcg {rep.pcg_iterations_total}, displacement vs oracle {err:.2e}")
        assert st == _lib.BH_OK and rep.converged
        assert rep.pcg_fallbacks == rep.iterations and rep.pcg_failures == 0
        assert err < (1e-8 if precision == 0 else 1e-5)
        st, q2, rep2, _ = _newton(mesh, 0, p.params, q0, fixed, f, max_it=2, gtol=1e-7, relative=1, pcg_tol=1e-10,
                                  precond=3, handle=P, precision=precision, fallback=0)
        print(f"  without the fallback: status {st}, failures {rep2.pcg_failures}, "
              f"|g|/|g0| {rep2.gradient_norm / rep2.gradient_norm0:.2e}")
        assert st == _lib.BH_MAXIT and not rep2.converged
        assert rep2.pcg_fallbacks == 0 and rep2.pcg_failures == rep2.iterations
        assert np.array_equal(q2, q0)  # zero directions: nothing moved (and nothing turned NaN)
    finally:
        lib.bh_precond_destroy(P)
        mesh.close()

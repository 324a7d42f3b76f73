"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle and the reference fixtures.

Tolerances (FP64 mode):
  * pattern, row pointers, column indices: bit-exact;
  * Hessian values: |dH| / |H| <= 1e-12 (norm-wise) -- element blocks are well conditioned;
  * energy and gradient at strained states: <= 1e-12 relative;
  * energy and gradient within 1e-3 h of rest (configuration 1): the energy density
    C (I1 - 3 - 2 ln J) cancels to O(strain^2), so any two correct FP64 evaluations differ at
    ~1e-16 * (term magnitude) / psi; the bound used is 1e-13 of the summed term magnitudes;
  * PSD-projected Hessians: 1e-10 (two different eigen-solvers, both FP64).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torch

    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from bartels_amd import tetmesh

    return tetmesh


def make_mesh(tm, V, E):
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    return tm.TetMesh(V.shape[0], E, dX, vol), dX, vol


def nrel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def csr_from_mesh(mesh, vals):
    rowptr, colind = mesh.pattern()
    n = 3 * mesh.nV
    return sp.csr_matrix((vals, colind, rowptr), shape=(n, n))


def rotation(rng):
    Q, R = np.linalg.qr(rng.standard_normal((3, 3)))
    Q = Q * np.sign(np.diag(R))
    return Q if np.linalg.det(Q) > 0 else -Q


def strained(V, seed, spread=0.1, wave=0.004):
    """smooth deformation: a stretch with singular values 1+spread, 1, 1-spread plus a transverse
    wave whose strain stays below ~2*pi*wave, so every tet keeps singular-value gaps well above the
    reference's 1e-5 perturbation threshold (src/d2psi_stretch_dF2.cpp:25-28)"""
    rng = np.random.default_rng(seed)
    A = rotation(rng) @ np.diag([1 + spread, 1.0, 1 - spread]) @ rotation(rng)
    L = V.max(0) - V.min(0) + 1e-12
    w = wave * L.max() * np.sin(2 * np.pi * V / L.max() + rng.random(3))
    return (V @ A.T + w[:, [1, 2, 0]]).ravel()


# ---- reference fixtures --------------------------------------------------------------------------
def test_bunny_pattern_and_legacy_hessian(tm, bunny):
    V, E = bunny["V"], bunny["E"]
    mesh, dX, vol = make_mesh(tm, V, E)
    rowptr, colind = mesh.pattern()
    assert np.array_equal(rowptr, bunny["indptr"]) and np.array_equal(colind, bunny["indices"])
    q = (V * np.array([2.0, 4.0, 3.0])).ravel()
    YM, nu = 6e5, 0.4
    D = 0.5 * (YM * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    C = 0.5 * YM / (2.0 * (1.0 + nu))
    params = np.tile([C, D], (E.shape[0], 1))
    vals = mesh.dq2_values(tm.NEOHOOKEAN_LEGACY, q, params)
    err = np.linalg.norm(vals - bunny["H"]) / vals[0]
    print("bunny legacy-NH Hessian vs H_neohookean_coarser_bunny.txt:", err)
    assert err < 1e-12


def test_bunny_mass_matrix_and_gravity(tm, bunny):
    import torch

    V, E = bunny["V"], bunny["E"]
    mesh, dX, vol = make_mesh(tm, V, E)
    dens = torch.full((E.shape[0],), 0.1, dtype=torch.float64, device="cuda")
    Mv = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
    mesh.mass_matrix_device(dens, Mv)
    g = np.array([0.0, -9.8, 0.0])
    f = torch.empty(3 * mesh.nV, dtype=torch.float64, device="cuda")
    mesh.const_ext_force_device(0.1 * g, f)
    torch.cuda.synchronize()
    Mv = Mv.cpu().numpy()
    assert np.linalg.norm(Mv - bunny["M"]) < 1e-8
    assert np.linalg.norm(Mv - bunny["M"]) / bunny["M"][0] < 1e-12
    M = csr_from_mesh(mesh, Mv)
    assert np.linalg.norm(f.cpu().numpy() - M @ np.tile(g, mesh.nV)) < 1e-8


def test_dphidX_host_and_device_match_oracle(tm, oracle):
    import torch

    V, E = meshgen.beam_grid(9, 5, 4, h=0.03)
    V = meshgen.jitter(V, E, 0.03, 0.2, seed=11)
    mesh, dX, vol = make_mesh(tm, V, E)
    dXo, volo = oracle.dphidX_volume(V, E)
    assert nrel(dX, dXo) < 1e-14 and nrel(vol, volo) < 1e-14
    Vd = torch.tensor(np.asfortranarray(V).ravel(order="F"), device="cuda")
    dXd = torch.empty(E.shape[0] * 12, dtype=torch.float64, device="cuda")
    vold = torch.empty(E.shape[0], dtype=torch.float64, device="cuda")
    mesh.dphi_dX_device(Vd, dXd, vold)
    torch.cuda.synchronize()
    assert nrel(dXd.cpu().numpy().reshape(12, -1).T, dXo) < 1e-14
    assert nrel(vold.cpu().numpy(), volo) < 1e-14


# ---- configuration 1 (Neo-Hookean, near rest) -----------------------------------------------------
def energy_term_scale(V, E, q, dX, vol, params):
    """sum_e vol (C (I1 + 3 + 2|ln J|) + D ((J-1)^2 + 1)): the magnitudes the NH energy cancels"""
    X = q.reshape(-1, 3)[E]  # nT x 4 x 3
    phi = np.asarray(dX).reshape(dX.shape[0], 3, 4).transpose(0, 2, 1)  # phi[e, m, j] = dphidX(e, m + 4j)
    F = np.einsum("tmi,tmj->tij", X, phi)
    J = np.linalg.det(F)
    I1 = (F * F).sum((1, 2))
    return float(np.sum(vol * (np.abs(params[:, 0]) * (I1 + 3 + 2 * np.abs(np.log(np.abs(J)))) +
                               np.abs(params[:, 1]) * ((J - 1) ** 2 + 1))))


def test_config1_parity(tm, oracle):
    p = meshgen.config1()
    mesh, dX, vol = make_mesh(tm, p.V, p.E)
    nV = p.V.shape[0]
    rowptr, colind = mesh.pattern()
    H_or = oracle.tetmesh_dq2(0, nV, p.E, p.q, dX, vol, p.params)
    assert np.array_equal(rowptr, H_or.indptr) and np.array_equal(colind, H_or.indices)
    assert mesh.nnz == 477342
    for name, q in (("config1", p.q), ("strained", strained(p.V, 5))):
        e_gpu = mesh.q(tm.NEOHOOKEAN, q, p.params)
        e_or = oracle.tetmesh_q(0, nV, p.E, q, dX, vol, p.params)
        g_gpu = mesh.dq(tm.NEOHOOKEAN, q, p.params)
        g_or = oracle.tetmesh_dq(0, nV, p.E, q, dX, vol, p.params)
        H_or = oracle.tetmesh_dq2(0, nV, p.E, q, dX, vol, p.params)
        v_gpu = mesh.dq2_values(tm.NEOHOOKEAN, q, p.params)
        H_gpu = csr_from_mesh(mesh, v_gpu)
        eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
        eE = abs(e_gpu - e_or) / abs(e_or)
        eG = nrel(g_gpu, g_or)
        scale = energy_term_scale(p.V, p.E, q, dX, vol, p.params)
        print(f"{name}: energy rel {eE:.2e} (vs term scale {abs(e_gpu - e_or) / scale:.2e}), "
              f"gradient rel {eG:.2e}, Hessian rel {eH:.2e}")
        assert eH < 1e-12
        if name == "strained":
            assert eE < 1e-12 and eG < 1e-12
        else:
            assert abs(e_gpu - e_or) < 1e-13 * scale
            # near rest the gradient is small next to its terms (measured 3.2e-13 relative)
            assert eG < 1e-12


# ---- every model, reference and textbook layouts, mixed ids ---------------------------------------
@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("layout", [1, 0])
def test_model_parity(tm, oracle, model, layout):
    if layout == 0 and model not in (2, 3):
        pytest.skip("layout only affects the stretch models")
    V, E = meshgen.beam_grid(12, 6, 5, h=0.02)
    q = strained(V, 100 + model)
    lam, mu = meshgen.emu_to_lame(np.linspace(1e5, 1e7, E.shape[0]), np.linspace(0.3, 0.48, E.shape[0]))
    params = meshgen.model_params(np.full(E.shape[0], model, np.uint8), lam, mu)
    mesh, dX, vol = make_mesh(tm, V, E)
    mesh.set_stretch_reference_layout(layout == 1)
    nV = V.shape[0]
    e_gpu, e_or = mesh.q(model, q, params), oracle.tetmesh_q(model, nV, E, q, dX, vol, params)
    g_gpu, g_or = mesh.dq(model, q, params), oracle.tetmesh_dq(model, nV, E, q, dX, vol, params)
    H_gpu = csr_from_mesh(mesh, mesh.dq2_values(model, q, params))
    H_or = oracle.tetmesh_dq2(model, nV, E, q, dX, vol, params, ref_layout=layout)
    eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
    print(f"model {model} layout {layout}: energy {abs(e_gpu - e_or) / abs(e_or):.2e} gradient {nrel(g_gpu, g_or):.2e} "
          f"Hessian {eH:.2e}")
    assert abs(e_gpu - e_or) / abs(e_or) < 1e-12
    assert nrel(g_gpu, g_or) < 1e-12
    assert eH < 1e-12


def test_generic_elastic_mixed(tm, oracle):
    p = meshgen.config3(grid=(20, 8, 8))
    mesh, dX, vol = make_mesh(tm, p.V, p.E)
    nV = p.V.shape[0]
    ids = p.model_ids
    assert set(np.unique(ids)) == {0, 1, 2}
    e_gpu = mesh.q(tm.MODEL_PER_TET, p.q, p.params, model_ids=ids)
    e_or = oracle.tetmesh_q(0, nV, p.E, p.q, dX, vol, p.params, model_ids=ids)
    g_gpu = mesh.dq(tm.MODEL_PER_TET, p.q, p.params, model_ids=ids)
    g_or = oracle.tetmesh_dq(0, nV, p.E, p.q, dX, vol, p.params, model_ids=ids)
    H_gpu = csr_from_mesh(mesh, mesh.dq2_values(tm.MODEL_PER_TET, p.q, p.params, model_ids=ids))
    H_or = oracle.tetmesh_dq2(0, nV, p.E, p.q, dX, vol, p.params, model_ids=ids)
    eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
    print(f"mixed: energy {abs(e_gpu - e_or) / abs(e_or):.2e} gradient {nrel(g_gpu, g_or):.2e} Hessian {eH:.2e}")
    assert abs(e_gpu - e_or) / abs(e_or) < 1e-12 and nrel(g_gpu, g_or) < 1e-12 and eH < 1e-12


def test_generic_elastic_mixed_psd(tm, oracle):
    """per-tet models (Neo-Hookean, StVK, corotational) with simple_psd_fix: the PSD pass dispatches
    each tet's reduced matrix on its model id; oracle parity at two tolerances"""
    p = meshgen.config3(grid=(20, 8, 8))
    mesh, dX, vol = make_mesh(tm, p.V, p.E)
    nV = p.V.shape[0]
    ids = p.model_ids
    q = p.q  # compressed toward J in [0.2, 1.5]: indefinite element Hessians
    for tol in (1e-3, 1e2):
        H_gpu = csr_from_mesh(mesh, mesh.dq2_values(tm.MODEL_PER_TET, q, p.params, callback=tm.simple_psd_fix(tol),
                                                    model_ids=ids))
        H_or = oracle.tetmesh_dq2(0, nV, p.E, q, dX, vol, p.params, model_ids=ids, psd=True, psd_tol=tol)
        eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
        print(f"mixed PSD tol {tol}: Hessian {eH:.2e}")
        assert eH < 1e-12


def oracle_rounding_floor(oracle, model, nV, E, q, dX, vol, params, draws=4):
    """max over a few draws of |H(q') - H(q)| / |H(q)| for the oracle, with q' = q (1 +- eps) (random
    signs): what FP64 rounding of the positions alone does to the Hessian.  It is large only where J
    is ill-conditioned (near-degenerate or inverted tets)."""
    H0 = oracle.tetmesh_dq2(model, nV, E, q, dX, vol, params)
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(draws):
        qp = q * (1.0 + np.finfo(np.float64).eps * rng.choice([-1.0, 1.0], q.shape))
        H1 = oracle.tetmesh_dq2(model, nV, E, qp, dX, vol, params)
        worst = max(worst, sp.linalg.norm(H1 - H0) / sp.linalg.norm(H0))
    return worst


@pytest.mark.parametrize("model", [0, 1, 2, 3, 4])
def test_psd_parity(tm, oracle, model):
    V, E = meshgen.beam_grid(10, 5, 5, h=0.02)
    rng = np.random.default_rng(42 + model)
    q = (V + 0.25 * 0.02 * rng.standard_normal(V.shape)).ravel()  # strongly distorted: many indefinite tets
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.full(E.shape[0], model, np.uint8), lam, mu)
    mesh, dX, vol = make_mesh(tm, V, E)
    nV = V.shape[0]
    # the element Hessians themselves, without the projection.  This distortion inverts and nearly
    # degenerates tets (|J| down to 5e-4), where the Neo-Hookean blocks (1/J^2, and J^-8/3 for the
    # isochoric model) amplify the rounding of J: there the bound is the oracle's own rounding floor,
    # the change of its Hessian when q is perturbed by one unit in the last place (1e-12 where that
    # floor is lower)
    H_raw = oracle.tetmesh_dq2(model, nV, E, q, dX, vol, params)
    e_raw = sp.linalg.norm(csr_from_mesh(mesh, mesh.dq2_values(model, q, params)) - H_raw) / sp.linalg.norm(H_raw)
    floor = oracle_rounding_floor(oracle, model, nV, E, q, dX, vol, params)
    print(f"model {model}: Hessian without PSD {e_raw:.2e}, oracle one-ulp floor {floor:.2e}")
    assert e_raw < max(1e-12, floor)
    for tol in (1e-3, 1e2, 3e4):  # 3e4: most eigenvalues below tol (the full-eigensolver pass)
        H_gpu = csr_from_mesh(mesh, mesh.dq2_values(model, q, params, callback=tm.simple_psd_fix(tol)))
        H_or = oracle.tetmesh_dq2(model, nV, E, q, dX, vol, params, psd=True, psd_tol=tol)
        eH = sp.linalg.norm(H_gpu - H_or) / sp.linalg.norm(H_or)
        changed = sp.linalg.norm(H_or - H_raw) / sp.linalg.norm(H_raw)
        print(f"PSD model {model} tol {tol}: rel {eH:.2e}, without PSD {e_raw:.2e} (projection changed H by "
              f"{changed:.2e})")
        assert changed > 1e-6  # the test exercises clamping
        # the projection adds no error beyond the element Hessian's own (and 1e-12 where that is small)
        assert eH < max(1e-12, 1.5 * e_raw)


def test_determinism(tm):
    import torch

    p = meshgen.config1(grid=(24, 8, 8))
    mesh, dX, vol = make_mesh(tm, p.V, p.E)
    q = torch.tensor(strained(p.V, 3), device="cuda")
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device="cuda")
    outs = []
    for _ in range(2):
        H = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
        g = torch.empty(3 * mesh.nV, dtype=torch.float64, device="cuda")
        e = torch.empty(1, dtype=torch.float64, device="cuda")
        mesh.dq2_device(0, q, params, H, callback=tm.simple_psd_fix(1e-3))
        mesh.dq_device(0, q, params, g)
        mesh.q_device(0, q, params, e)
        torch.cuda.synchronize()
        outs.append((H.cpu().numpy(), g.cpu().numpy(), e.cpu().numpy()))
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("psd", [False, True])
def test_assembled_hessian_exactly_symmetric(tm, fp32, psd):
    """with the PSD projection the assembled matrix is symmetric to the last bit, FP64 and FP32:
    block (n, v) is the exact transpose of block (v, n) (the precision form's implied diagonal and
    the split operator both rely on it); the element-math gathers without it are symmetric to
    rounding"""
    import torch

    p = meshgen.config1(grid=(24, 8, 8))
    mesh, dX, vol = make_mesh(tm, p.V, p.E)
    q = torch.tensor(strained(p.V, 5), device="cuda")
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device="cuda")
    if fp32:
        params = params.to(torch.float32)
    import ctypes

    from bartels_amd import _lib

    lib = _lib.load()
    for lanes in (1, 2, 4, 8):
        assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(lanes, 0))) == 0
        H = torch.empty(mesh.nnz, dtype=params.dtype, device="cuda")
        fn = mesh.dq2_f32_device if fp32 else mesh.dq2_device
        fn(0, q, params, H, callback=tm.simple_psd_fix(1e-3) if psd else None)
        torch.cuda.synchronize()
        A = csr_from_mesh(mesh, H.cpu().numpy().astype(np.float64))
        D = (A - A.T).tocoo()
        asym = np.abs(D.data).max(initial=0.0) / np.abs(A.data).max()
        print(f"fp32={fp32} psd={psd} lanes={lanes}: max |H - H^T| / max |H| = {asym:.3e}")
        if psd:
            assert asym == 0.0
        else:
            assert asym < (1e-6 if fp32 else 1e-14)
    assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(4, 0))) == 0


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("psd", [True, False])
def test_gather_variants_agree(tm, fp32, psd):
    """every kernel schedule -- 1, 2, 4 or 8 lanes per vertex, XCD grouping on/off, vertex groups visited
    in vertex order or in the Morton order of their rest centroids -- keeps each slot's ascending-element
    fold and the same element math: same results to rounding across lane counts, bitwise equal for one
    lane count whatever the visiting order and XCD grouping, and bitwise equal for 4 and 8 lanes (the
    same block arithmetic; the two incidences an 8-lane step takes fold in list order)"""
    import ctypes

    import torch

    from bartels_amd import _lib

    p = meshgen.config1(grid=(24, 8, 8))
    dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
    mesh = tm.TetMesh(p.V.shape[0], p.E, dX, vol, V=p.V)  # rest positions: the Morton visiting order
    q = torch.tensor(strained(p.V, 4), device="cuda")
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device="cuda")
    if fp32:
        params = params.to(torch.float32)
    lib = _lib.load()
    outs = {}
    for lanes in (1, 2, 4, 8):
        for xcd in (0, 1):
            for order in (1, 2, 3):
                t = _lib.BhTuning(lanes, xcd, 0, 0, order)
                assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(t)) == 0
                H = torch.empty(mesh.nnz, dtype=params.dtype, device="cuda")
                fn = mesh.dq2_f32_device if fp32 else mesh.dq2_device
                fn(0, q, params, H, callback=tm.simple_psd_fix(1e-3) if psd else None)
                torch.cuda.synchronize()
                outs[(lanes, xcd, order)] = H.cpu().numpy().astype(np.float64)
    assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(4, 0, 0, 0, 3))) == 0
    ref = outs[(1, 0, 2)]
    worst = 0.0
    for k, v in outs.items():
        d = np.abs(v - ref).max() / np.abs(ref).max()
        worst = max(worst, d)
        print(f"schedule (lanes, xcd, order) {k}: max rel diff {d:.2e}, bitwise equal: {np.array_equal(v, ref)}")
        assert np.array_equal(v, outs[(k[0], 0, 2)]), f"schedule {k} differs from vertex order"
        if k[0] == 8:
            assert np.array_equal(v, outs[(4,) + k[1:]]), f"schedule {k} differs from four lanes"
    assert worst < (1e-6 if fp32 else 1e-14)


@pytest.mark.parametrize("psd", [False, True])
def test_gather_shapes_small_and_ragged(tm, oracle, psd):
    """every gather shape (1, 2, 4, 8 lanes per vertex) on small ragged meshes -- one tet; two tets
    sharing a face (odd vertex degrees, so an eight-lane step has one idle quad); a tet fan around an
    edge; a mesh with an unused vertex -- against the oracle, with and without the clamp"""
    import ctypes

    from bartels_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(5)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], float) * 0.1
    fan = np.array([[0, 0, 0], [0, 0, 0.1]] + [[0.1 * np.cos(t), 0.1 * np.sin(t), 0.05]
                                               for t in np.linspace(0, 2 * np.pi, 7)[:-1]])
    cases = {
        "one tet": (base[:4], np.array([[0, 1, 2, 3]])),
        "two tets sharing a face": (base, np.array([[0, 1, 2, 3], [1, 2, 3, 4]])),
        "fan of six tets around an edge": (fan, np.array([[0, 1, 2 + k, 2 + (k + 1) % 6] for k in range(6)])),
        "unused vertex": (np.vstack([base, [[0.5, 0.5, 0.5]]]), np.array([[0, 1, 2, 3], [1, 2, 3, 4]])),
    }
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    for name, (V, E) in cases.items():
        E = np.array(E, dtype=np.int32)
        d = np.einsum("ij,ij->i", np.cross(V[E[:, 1]] - V[E[:, 0]], V[E[:, 2]] - V[E[:, 0]]), V[E[:, 3]] - V[E[:, 0]])
        E[d < 0] = E[d < 0][:, [0, 1, 3, 2]]  # positively oriented
        mesh, dX, vol = make_mesh(tm, V, E)
        q = (V + 0.02 * rng.standard_normal(V.shape)).ravel()
        params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
        H_or = oracle.tetmesh_dq2(0, V.shape[0], E, q, dX, vol, params, psd=psd, psd_tol=1e-3)
        outs = {}
        for lanes in (1, 2, 4, 8):
            assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(lanes, 0))) == 0
            H = csr_from_mesh(mesh, mesh.dq2_values(0, q, params, callback=tm.simple_psd_fix(1e-3) if psd else None))
            outs[lanes] = H
            e = sp.linalg.norm(H - H_or) / sp.linalg.norm(H_or)
            print(f"{name}, psd={psd}, lanes {lanes}: {e:.2e}")
            assert e < 1e-12, (name, lanes)
        assert (outs[8] != outs[4]).nnz == 0, name  # four and eight lanes: the same bits

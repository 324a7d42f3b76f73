"""GPU parity of the triangle siblings (tri2d plane-strain elements, the tri3d StVK shell element)
through the C ABI against the oracle: pattern bit-exact; energy, gradient and Hessian (with and
without simple_psd_fix) to 1e-12 relative in FP64; the mass matrix; the shell Hessian with per-element
normals and normal derivatives."""
import numpy as np
import pytest
import scipy.sparse as sp

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu

MODELS = [0, 1, 2, 3]


@pytest.fixture(scope="module")
def tri():
    import torch

    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from bartels_amd import trimesh

    return trimesh


def deformed_strip(seed):
    V, E = meshgen.strip2d(20, 6, h=0.05, seed=seed, jitter_amount=0.15)
    rng = np.random.default_rng(seed)
    A = np.array([[1.15, 0.25], [-0.1, 0.85]])
    q = (V @ A.T).ravel() + 0.002 * rng.standard_normal(V.size)
    return V, E, q


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.mark.parametrize("model", MODELS)
def test_tri2d_parity(tri, oracle, model):
    V, E, q = deformed_strip(20 + model)
    nV, nT = V.shape[0], E.shape[0]
    dX = tri.linear_tri2dmesh_dphi_dX(V, E)
    assert np.abs(dX - oracle.tri2d_dphidX(V, E)).max() < 1e-12 * np.abs(dX).max()
    vol = meshgen.triangle_areas(V, E)
    lam, mu = meshgen.emu_to_lame(np.linspace(1e5, 1e7, nT), np.linspace(0.3, 0.48, nT))
    params = meshgen.model_params(np.full(nT, model, np.uint8), lam, mu)
    mesh = tri.TriMesh(2, nV, E)
    rowptr, colind = mesh.pattern()
    H_or = oracle.tri2dmesh_dq2(model, nV, E, q, dX, vol, params)
    assert np.array_equal(rowptr, H_or.indptr) and np.array_equal(colind, H_or.indices)
    e_g, e_o = mesh.q(model, q, dX, vol, params), oracle.tri2dmesh_q(model, nV, E, q, dX, vol, params)
    g_g, g_o = mesh.dq(model, q, dX, vol, params), oracle.tri2dmesh_dq(model, nV, E, q, dX, vol, params)
    H_g = mesh.dq2(model, q, dX, vol, params)
    Hp_g = mesh.dq2(model, q, dX, vol, params, callback=tri.simple_psd_fix(1e-3))
    Hp_o = oracle.tri2dmesh_dq2(model, nV, E, q, dX, vol, params, psd=True, psd_tol=1e-3)
    eE, eG = abs(e_g - e_o) / abs(e_o), rel(g_g, g_o)
    eH = sp.linalg.norm(H_g - H_or) / sp.linalg.norm(H_or)
    eP = sp.linalg.norm(Hp_g - Hp_o) / sp.linalg.norm(Hp_o)
    print(f"tri2d model {model}: energy {eE:.2e}, gradient {eG:.2e}, Hessian {eH:.2e}, PSD Hessian {eP:.2e}")
    assert eE < 1e-12 and eG < 1e-12 and eH < 1e-12 and eP < 1e-11
    # the wrapper functions with the reference's names give the same matrix
    H2 = getattr(tri, ["linear_tri2dmesh_neohookean_dq2", "linear_tri2dmesh_stvk_dq2",
                       "linear_tri2dmesh_corotational_dq2", "linear_tri2dmesh_arap_dq2"][model])(V, E, q, dX, vol, params)
    assert np.array_equal(H2.data, H_g.data)
    mesh.close()


def test_tri2d_mass_matrix(tri, oracle):
    V, E, _ = deformed_strip(7)
    vol = meshgen.triangle_areas(V, E)
    rho = np.linspace(900.0, 1100.0, E.shape[0])
    M_g = tri.linear_tri2dmesh_mass_matrix(V, E, rho, vol)
    M_o = oracle.tri2dmesh_mass(V.shape[0], E, rho, vol)
    assert np.array_equal(M_g.indptr, M_o.indptr) and np.array_equal(M_g.indices, M_o.indices)
    assert np.abs((M_g - M_o).data).max() < 1e-13 * np.abs(M_o.data).max()


@pytest.mark.parametrize("psd", [False, True])
def test_tri3d_stvk_shell_parity(tri, oracle, psd):
    rng = np.random.default_rng(31)
    V2, E = meshgen.strip2d(12, 5, h=0.05, seed=31, jitter_amount=0.1)
    V = np.column_stack([V2, 0.02 * np.sin(V2[:, 0] * 9.0)])  # a gently curved sheet
    nV, nT = V.shape[0], E.shape[0]
    dX = tri.linear_tri3dmesh_dphi_dX(V, E)
    assert np.abs(dX - oracle.tri3d_dphidX(V, E)).max() < 1e-12 * np.abs(dX).max()
    N = np.cross(V[E[:, 1]] - V[E[:, 0]], V[E[:, 2]] - V[E[:, 0]])
    N /= np.linalg.norm(N, axis=1)[:, None]
    q = (V + 0.003 * rng.standard_normal(V.shape)).ravel()
    x = q.reshape(-1, 3)
    n = np.cross(x[E[:, 1]] - x[E[:, 0]], x[E[:, 2]] - x[E[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    dndq = 0.05 * rng.standard_normal((nT, 27))
    vol = 0.5 * np.linalg.norm(np.cross(V[E[:, 1]] - V[E[:, 0]], V[E[:, 2]] - V[E[:, 0]]), axis=1) * 1e-3
    lam, mu = meshgen.emu_to_lame(np.full(nT, 1e6), np.full(nT, 0.3))
    params = np.stack([lam, mu], 1)
    cb = tri.simple_psd_fix(1e-3) if psd else None
    H_g = tri.linear_tri3dmesh_stvk_dq2(V, E, q, dX, N, n, dndq, vol, params, cb)
    H_o = oracle.tri3dmesh_stvk_dq2(nV, E, q, dX, N, n, dndq, vol, params, psd=psd, psd_tol=1e-3)
    assert np.array_equal(H_g.indptr, H_o.indptr) and np.array_equal(H_g.indices, H_o.indices)
    e = sp.linalg.norm(H_g - H_o) / sp.linalg.norm(H_o)
    print(f"tri3d StVK shell (psd={psd}): Hessian {e:.2e}")
    assert e < (1e-11 if psd else 1e-12)


@pytest.mark.parametrize("tol", [1e-3, 1e2, 0.0, -1e-3])
def test_tri3d_shell_clamp_consistent_normals(tri, oracle, tol):
    """the shell clamp with the geometric normal derivatives (they sum to zero over the three vertices,
    as the shape gradients do), so the element Hessians annihilate the translations and the clamp runs
    through the 6x6 reduced matrix; against the oracle's 9x9 eigensolver, at tolerances that clamp
    nothing, some or most of the spectrum, and that leave the translations at zero"""
    rng = np.random.default_rng(37)
    V2, E = meshgen.strip2d(12, 5, h=0.05, seed=37, jitter_amount=0.1)
    V = np.column_stack([V2, 0.02 * np.sin(V2[:, 0] * 9.0)])
    nV, nT = V.shape[0], E.shape[0]
    dX = tri.linear_tri3dmesh_dphi_dX(V, E)
    N = np.cross(V[E[:, 1]] - V[E[:, 0]], V[E[:, 2]] - V[E[:, 0]])
    N /= np.linalg.norm(N, axis=1)[:, None]
    x = V + 0.01 * rng.standard_normal(V.shape)
    x[:, 0] *= 0.97  # some compression: negative eigenvalues to clamp
    q = x.ravel()
    e1, e2 = x[E[:, 1]] - x[E[:, 0]], x[E[:, 2]] - x[E[:, 0]]
    c = np.cross(e1, e2)
    cn = np.linalg.norm(c, axis=1)
    n = c / cn[:, None]
    Pn = (np.eye(3)[None] - n[:, :, None] * n[:, None, :]) / cn[:, None, None]

    def skew(v):
        z = np.zeros(v.shape[0])
        return np.stack([np.stack([z, -v[:, 2], v[:, 1]], 1), np.stack([v[:, 2], z, -v[:, 0]], 1),
                         np.stack([-v[:, 1], v[:, 0], z], 1)], 1)

    dc1, dc2 = -skew(e2), skew(e1)
    dndq = np.concatenate([Pn @ d_ for d_ in (-(dc1 + dc2), dc1, dc2)], axis=2).reshape(nT, 27, order="F")
    vol = 0.5 * np.linalg.norm(np.cross(V[E[:, 1]] - V[E[:, 0]], V[E[:, 2]] - V[E[:, 0]]), axis=1) * 1e-3
    lam, mu = meshgen.emu_to_lame(np.full(nT, 1e6), np.full(nT, 0.3))
    params = np.stack([lam, mu], 1)
    H_g = tri.linear_tri3dmesh_stvk_dq2(V, E, q, dX, N, n, dndq, vol, params, tri.simple_psd_fix(tol))
    H_o = oracle.tri3dmesh_stvk_dq2(nV, E, q, dX, N, n, dndq, vol, params, psd=True, psd_tol=tol)
    H_raw = oracle.tri3dmesh_stvk_dq2(nV, E, q, dX, N, n, dndq, vol, params, psd=False)
    e = sp.linalg.norm(H_g - H_o) / sp.linalg.norm(H_o)
    moved = sp.linalg.norm(H_o - H_raw) / sp.linalg.norm(H_raw)
    print(f"tri3d shell clamp, consistent normals, tol {tol}: Hessian {e:.2e} (the clamp moved H by {moved:.2e})")
    assert e < 1e-11

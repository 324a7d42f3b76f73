"""Full-size (configuration 2: 64x64x256 beam, 6,072,570 tets, 138,917,358 nonzeros) checks of the
GPU path through size-independent properties, plus oracle parity on sampled rows: the block row of
a vertex depends only on the tets incident to it, so the oracle run on just those tets must
reproduce the sampled rows of the full-size GPU Hessian."""
import numpy as np
import pytest

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    import torch

    assert torch.cuda.is_available()
    from bartels_amd import tetmesh as tm

    p = meshgen.config2()
    dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
    mesh = tm.TetMesh(p.V.shape[0], p.E, dX, vol)
    q = torch.tensor(p.q, device="cuda")
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device="cuda")
    return dict(p=p, dX=dX, vol=vol, mesh=mesh, q=q, params=params, tm=tm)


def test_config2_sizes(big):
    assert big["mesh"].nT == 6072570 and big["mesh"].nnz == 138917358


def test_config2_sampled_rows_match_oracle(big, oracle):
    import torch

    p, mesh, tm = big["p"], big["mesh"], big["tm"]
    H = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
    for psd in (None, tm.simple_psd_fix(1e-3)):
        mesh.dq2_device(0, big["q"], big["params"], H, callback=psd)
        torch.cuda.synchronize()
        vals = H.cpu().numpy()
        rowptr, colind = mesh.pattern()
        rng = np.random.default_rng(2)
        nV = p.V.shape[0]
        verts = np.unique(np.concatenate([rng.integers(0, nV, 40), [0, nV - 1, nV // 2]]))
        tets = np.nonzero(np.isin(p.E, verts).any(axis=1))[0]
        Eo = p.E[tets]
        Ho = oracle.tetmesh_dq2(0, nV, Eo, p.q, big["dX"][tets], big["vol"][tets], p.params[tets],
                                psd=psd is not None, psd_tol=1e-3).tocsr()
        worst = 0.0
        for v in verts:
            for l in range(3):
                r = 3 * v + l
                a, b = rowptr[r], rowptr[r + 1]
                assert np.array_equal(colind[a:b], Ho.indices[Ho.indptr[r]:Ho.indptr[r + 1]])
                ref = Ho.data[Ho.indptr[r]:Ho.indptr[r + 1]]
                worst = max(worst, np.linalg.norm(vals[a:b] - ref) / np.linalg.norm(ref))
        print(f"config2 sampled rows (psd={psd is not None}): worst row rel {worst:.2e} over {verts.size} vertices")
        assert worst < (1e-10 if psd is not None else 1e-12)


def test_config2_invariants(big):
    import torch

    p, mesh = big["p"], big["mesh"]
    n = 3 * mesh.nV
    q, params = big["q"], big["params"]
    g = torch.empty(n, dtype=torch.float64, device="cuda")
    mesh.dq_device(0, q, params, g)
    # translation invariance of the gradient
    gs = g.view(-1, 3).sum(0).abs().max().item()
    assert gs < 1e-9 * g.abs().sum().item(), gs
    # directional derivative of the energy matches g.d
    d = torch.tensor(meshgen.normal(9, n), device="cuda") * 1e-3
    e = torch.empty(2, dtype=torch.float64, device="cuda")
    h = 1e-4
    mesh.q_device(0, q + h * d, params, e[0:1])
    mesh.q_device(0, q - h * d, params, e[1:2])
    fd = (e[0] - e[1]).item() / (2 * h)
    gd = torch.dot(g, d).item()
    print(f"config2 energy FD {fd:.10e} vs g.d {gd:.10e}")
    assert abs(fd - gd) < 1e-6 * abs(gd)
    # Hessian-vector product matches the gradient difference
    H = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
    mesh.dq2_device(0, q, params, H)
    Hd = torch.empty(n, dtype=torch.float64, device="cuda")
    mesh.spmv_device(H, d, Hd)
    gp, gm = torch.empty_like(g), torch.empty_like(g)
    mesh.dq_device(0, q + h * d, params, gp)
    mesh.dq_device(0, q - h * d, params, gm)
    fdH = (gp - gm) / (2 * h)
    rel = (torch.linalg.norm(fdH - Hd) / torch.linalg.norm(Hd)).item()
    print(f"config2 Hessian-vector vs gradient FD: {rel:.2e}")
    assert rel < 1e-6
    # the unprojected Hessian annihilates translations
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    t[0::3] = 1.0
    mesh.spmv_device(H, t, Hd)
    assert (Hd.abs().max() / H.abs().max()).item() < 1e-9
    # the PSD-projected one maps them to tol * (incident tets) * t
    mesh.dq2_device(0, q, params, H, callback=big["tm"].simple_psd_fix(1e-3))
    mesh.spmv_device(H, t, Hd)
    cnt = np.bincount(p.E.ravel(), minlength=mesh.nV).astype(np.float64)
    expect = torch.zeros(n, dtype=torch.float64, device="cuda")
    expect[0::3] = torch.tensor(1e-3 * cnt, device="cuda")
    assert ((Hd - expect).abs().max() / H.abs().max()).item() < 1e-9

"""On-device mesh set-up (bh_mesh_create_device) against the host build (bh_mesh_create): the block
pattern, the CSR pattern and every energy / gradient / Hessian output identical to the last bit (the
incidence lists and ranks the gathers read must then be the same too), on beams, on a randomly
renumbered beam with an unused vertex, and at the 64x64x256 size; with dphidX and the signed volume
computed on the device from V as well (within rounding of the host's)."""
import ctypes
import time

import numpy as np
import pytest

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


def _block_pattern(mesh):
    import torch

    from bartels_amd import _lib

    lib = _lib.load()
    bp, bc = ctypes.c_void_p(), ctypes.c_void_p()
    nb = ctypes.c_int64()
    assert lib.bh_mesh_block_pattern_device(mesh.handle, ctypes.byref(bp), ctypes.byref(bc), ctypes.byref(nb)) == 0
    out_p = np.zeros(mesh.nV + 1, np.int32)
    out_c = np.zeros(max(1, nb.value), np.int32)
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library is linked against
    assert hip.hipMemcpy(out_p.ctypes.data_as(ctypes.c_void_p), bp, ctypes.c_size_t(4 * (mesh.nV + 1)), 2) == 0
    if nb.value:
        assert hip.hipMemcpy(out_c.ctypes.data_as(ctypes.c_void_p), bc, ctypes.c_size_t(4 * nb.value), 2) == 0
    return out_p, out_c[:nb.value]


def _outputs(mesh, q, params):
    import torch

    from bartels_amd import tetmesh as tm

    H = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
    Hp = torch.empty(mesh.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(3 * mesh.nV, dtype=torch.float64, device="cuda")
    e = torch.empty(1, dtype=torch.float64, device="cuda")
    mesh.dq2_device(0, q, params, H)
    mesh.dq2_device(0, q, params, Hp, callback=tm.simple_psd_fix(1e-3))
    mesh.dq_device(0, q, params, g)
    mesh.q_device(0, q, params, e)
    torch.cuda.synchronize()
    return H, Hp, g, e


def _compare(V, E, q):
    import torch

    from bartels_amd import tetmesh as tm

    nV, nT = V.shape[0], E.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.zeros(nT, np.uint8), lam, mu)
    t0 = time.perf_counter()
    host = tm.TetMesh(nV, E, dX, vol)
    t_host = time.perf_counter() - t0
    dE = torch.tensor(np.asfortranarray(E).ravel(order="F"), dtype=torch.int32, device="cuda")
    ddX = torch.tensor(np.asfortranarray(dX).ravel(order="F"), device="cuda")
    dvol = torch.tensor(vol, device="cuda")
    dV = torch.tensor(np.asfortranarray(V).ravel(order="F"), device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = tm.TetMesh.from_device(nV, dE, dphidX=ddX, volume=dvol)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    devV = tm.TetMesh.from_device(nV, dE, V=dV)
    t_devV = time.perf_counter() - t0
    try:
        assert dev.nnz == host.nnz == devV.nnz
        bp_h, bc_h = _block_pattern(host)
        bp_d, bc_d = _block_pattern(dev)
        assert np.array_equal(bp_h, bp_d) and np.array_equal(bc_h, bc_d)
        rp_h, ci_h = host.pattern()
        rp_d, ci_d = dev.pattern()
        assert np.array_equal(rp_h, rp_d) and np.array_equal(ci_h, ci_d)
        dq = torch.tensor(q, device="cuda")
        dpar = torch.tensor(np.asfortranarray(params).ravel(order="F"), device="cuda")
        for a, b in zip(_outputs(host, dq, dpar), _outputs(dev, dq, dpar)):
            assert torch.isfinite(a).all()  # no inverted tet in the deformation
            assert torch.equal(a.view(torch.int64), b.view(torch.int64))
        # dphidX from V on the device: the shape gradients differ from the host's by rounding (FMA
        # contraction), which the sums over incident tets amplify where they cancel
        for name, a, b in zip(("H", "H psd", "g", "energy"), _outputs(host, dq, dpar), _outputs(devV, dq, dpar)):
            rel = (torch.linalg.norm(a - b) / torch.linalg.norm(a)).item()
            print(f"  {name}: host dphidX vs device dphidX {rel:.2e}")
            assert rel < 1e-10
        return t_host, t_dev, t_devV
    finally:
        host.close()
        dev.close()
        devV.close()


@pytest.mark.parametrize("grid", [(40, 10, 10), (17, 6, 9)])
def test_device_setup_matches_host(grid):
    V, E = meshgen.beam_grid(*grid)
    V = meshgen.jitter(V, E, meshgen.H_DEFAULT, 0.2, seed=3)
    q = (V + 0.02 * meshgen.H_DEFAULT * meshgen.normal(2, V.size).reshape(V.shape)).ravel()
    _compare(V, E, q)


def test_device_setup_scattered_numbering():
    """vertices renumbered at random, tets shuffled, and one vertex used by no tet"""
    V, E = meshgen.beam_grid(12, 5, 5)
    rng = np.random.default_rng(4)
    nV = V.shape[0]
    perm = rng.permutation(nV + 1)  # new vertex i is old vertex perm[i]; old vertex nV is unused
    V1 = np.vstack([V, [[9.0, 9.0, 9.0]]])
    inv = np.empty(nV + 1, np.int64)
    inv[perm] = np.arange(nV + 1)
    E2 = inv[E].astype(np.int32)[rng.permutation(E.shape[0])]
    V2 = V1[perm]
    q = (V2 + 0.02 * meshgen.H_DEFAULT * rng.standard_normal(V2.shape)).ravel()
    _compare(V2, E2, q)


def test_device_setup_at_scale():
    """the 64x64x256 beam: identical arrays and outputs, and the set-up times (host vs device)"""
    p = meshgen.config2()
    t_host, t_dev, t_devV = _compare(p.V, p.E, p.q)
    print(f"64x64x256 beam ({p.E.shape[0]} tets): bh_mesh_create {t_host:.3f} s, bh_mesh_create_device "
          f"{t_dev:.3f} s, with dphidX and volume from V {t_devV:.3f} s")


def test_trim_then_recompute_is_identical():
    """bh_mesh_trim frees the buffers a mesh keeps between calls; every output made after it (the PSD
    records and Q factors, the corner forces and the FP32 copies are made again) equals the output
    made before it to the last bit, and trimming twice or before any call is harmless"""
    import torch

    from bartels_amd import tetmesh as tm

    V, E = meshgen.beam_grid(16, 6, 6, 0.05)
    V = meshgen.jitter(V, E, 0.05, 0.1, 3)
    nV = V.shape[0]
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    mesh = tm.TetMesh(nV, E, dX, vol, V=V)
    try:
        mesh.trim()  # nothing made yet
        lam, mu = meshgen.emu_to_lame(2e6, 0.45)
        params = torch.tensor(np.asfortranarray(meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu))
                              .ravel(order="F"), device="cuda")
        rng = np.random.default_rng(5)
        q = torch.tensor(V.ravel() + 0.001 * rng.standard_normal(3 * nV), device="cuda")
        before = _outputs(mesh, q, params)
        assert all(bool(torch.isfinite(t).all()) for t in before)
        H32 = torch.empty(mesh.nnz, dtype=torch.float32, device="cuda")
        mesh.dq2_f32_device(0, q, params.to(torch.float32), H32, callback=tm.simple_psd_fix(1e-3))
        torch.cuda.synchronize()
        h32 = H32.clone()
        mesh.trim()
        mesh.trim()
        after = _outputs(mesh, q, params)
        mesh.dq2_f32_device(0, q, params.to(torch.float32), H32, callback=tm.simple_psd_fix(1e-3))
        torch.cuda.synchronize()
        for name, a, b in zip(("dq2", "dq2 + PSD", "dq", "q"), before, after):
            print(f"{name}: identical after trim: {bool(torch.equal(a, b))}")
            assert torch.equal(a, b)
        assert torch.equal(h32, H32)
    finally:
        mesh.close()

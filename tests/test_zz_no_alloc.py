"""The allocation-free energy entry point (bh_linear_tetmesh_q_device_noalloc) against the ordinary
one (bh_linear_tetmesh_q_device): the same kernels and summation order with the partial sums in the
library's fixed device array instead of the mesh's own buffer, so the energies must agree bit for
bit.  Two calls in a row, the second on a smaller mesh (fewer partials than the first left in the
array), for a single model and for per-tet model ids."""
import ctypes

import numpy as np
import pytest

from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


def _energies(mesh, model, q, params, ids):
    import torch

    from bartels_amd import _lib

    lib = _lib.load()
    dq = torch.tensor(q, device="cuda")
    dp = torch.tensor(np.asfortranarray(params).ravel(order="F"), device="cuda")
    dids = torch.tensor(ids, device="cuda") if ids is not None else None
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    idp = ctypes.c_void_p(dids.data_ptr()) if dids is not None else None
    st0 = lib.bh_linear_tetmesh_q_device(mesh.handle, model, idp, ctypes.c_void_p(dq.data_ptr()),
                                         ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(out.data_ptr()), s)
    st1 = lib.bh_linear_tetmesh_q_device_noalloc(mesh.handle, model, idp, ctypes.c_void_p(dq.data_ptr()),
                                                 ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(out.data_ptr() + 8),
                                                 s)
    assert st0 == _lib.BH_OK and st1 == _lib.BH_OK, (st0, st1)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("per_tet", [False, True])
def test_noalloc_energy_matches(oracle, per_tet):
    from bartels_amd import tetmesh as tm

    # first a larger mesh, then a smaller one: the second call's reduction must read only the
    # partials its own energy kernel wrote
    for grid in ((64, 16, 16), (12, 5, 4)):
        p = meshgen.config3(grid=grid) if per_tet else meshgen.config2(grid=grid)
        nV = p.V.shape[0]
        dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
        mesh = tm.TetMesh(nV, p.E, dX, vol)
        try:
            model = tm.MODEL_PER_TET if per_tet else 0
            ids = p.model_ids if per_tet else None
            e_ref, e_noalloc = _energies(mesh, model, p.q, p.params, ids)
            e_or = oracle.tetmesh_q(0, nV, p.E, p.q, dX, vol, p.params, model_ids=ids)
            print(f"grid {grid} ({p.E.shape[0]} tets): energy {e_ref:.17g}, allocation-free {e_noalloc:.17g}, "
                  f"oracle rel {abs(e_ref - e_or) / abs(e_or):.2e}")
            assert np.float64(e_ref).tobytes() == np.float64(e_noalloc).tobytes()
            assert abs(e_noalloc - e_or) < 1e-12 * abs(e_or)
        finally:
            mesh.close()

"""CPU checks of the boundary and the synthetic inputs (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from bartels_amd import meshgen
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "bartels_hip.h")
LIB = os.path.join(ROOT, "bartels_amd", "libbartels_hip.so")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[A-Za-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built (python -m bartels_amd.build)")
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(LIB)  # loading resolves symbols; no device call is made
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    from bartels_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_product_path_has_no_cpu_fallback(monkeypatch, tmp_path):
    from bartels_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        _lib.load()


def test_rng_is_portable():
    # splitmix64 reference values (seed 0): first outputs of the published generator
    z = meshgen.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = meshgen.uniform(5, 1000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.05
    n = meshgen.normal(5, 20000)
    assert abs(n.mean()) < 0.03 and abs(n.std() - 1.0) < 0.03


@pytest.mark.parametrize("grid,nT,nnz", [((40, 10, 10), 18954, 477342), ((96, 24, 24), 301530, 7096878)])
def test_beam_sizes(grid, nT, nnz):
    """tet and CSR sizes of the configurations (Kuhn split along the main diagonal)"""
    V, E = meshgen.beam_grid(*grid)
    assert E.shape[0] == nT
    assert (meshgen.signed_volumes(V, E) > 0).all()
    # nnz = 9 (nV + 2 * #edges)
    edges = set()
    for a in range(4):
        for b in range(a + 1, 4):
            u, v = np.minimum(E[:, a], E[:, b]), np.maximum(E[:, a], E[:, b])
            edges.update((u.astype(np.int64) << 32 | v).tolist())
    assert 9 * (V.shape[0] + 2 * len(edges)) == nnz


def test_config1_shapes():
    p = meshgen.config1()
    assert p.V.shape == (4000, 3) and p.E.shape == (18954, 4) and p.q.shape == (12000,)
    assert p.fixed.size == 100 and np.all(p.V[p.fixed, 0] == 0)
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    assert np.allclose(p.params[0], [mu / 2, lam / 2])
    assert np.abs(p.q - p.V.ravel()).max() < 6 * 1e-3 * 0.025


def test_jitter_keeps_volumes_positive():
    V, E = meshgen.beam_grid(12, 6, 6)
    Vj = meshgen.jitter(V, E, 0.025, 0.2, seed=3)
    assert (meshgen.signed_volumes(Vj, E) > 0).all()
    assert np.abs(Vj - V).max() <= 0.2 * 0.025 + 1e-15


def test_bunny_reader(bunny):
    V, E = bunny["V"], bunny["E"]
    assert V.shape == (500, 3) and E.shape == (1518, 4) and E.min() == 0 and E.max() == 499


def test_host_dphidX_matches_oracle(oracle):
    """host-side setup of the library (no device call): dphidX and signed volume"""
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    from bartels_amd import tetmesh as tm

    V, E = meshgen.beam_grid(7, 5, 4, h=0.03)
    V = meshgen.jitter(V, E, 0.03, 0.2, seed=4)
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    dXo, volo = oracle.dphidX_volume(V, E)
    assert np.abs(dX - dXo).max() <= 1e-13 * np.abs(dXo).max()
    assert np.abs(vol - volo).max() <= 1e-15 * np.abs(volo).max()
    # F = diag(2,4,3) known answer through the library's dphidX
    q = (V * np.array([2.0, 4.0, 3.0]))
    phi = dX.reshape(-1, 3, 4).transpose(0, 2, 1)
    F = np.einsum("tmi,tmj->tij", q[E], phi)
    assert np.abs(F - np.diag([2.0, 4.0, 3.0])).max() < 1e-12

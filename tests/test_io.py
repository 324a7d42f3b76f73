"""CPU tests of the file formats (bh_read_tetgen, bh_read_medit, bh_save_market, bh_load_market): no GPU
call.  The bunny fixture (tests/golden/bunny.npz) keeps the SHA-256 digests of the reference's MEDIT mesh
and MatrixMarket files: writing the fixture's matrices back out must reproduce those files byte for byte
(Eigen's saveMarket format), and reading the reference's files (when /root/reference is mounted) must give
the fixture's arrays."""
import hashlib
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import REFERENCE

REF_DATA = os.path.join(REFERENCE, "data")


def _bunny_csc(bunny, key):
    return sp.csc_matrix((bunny[key], bunny["indices"], bunny["indptr"]), shape=(1500, 1500))


@pytest.mark.parametrize("key,digest_index", [("M", 1), ("H", 2)])
def test_save_market_reproduces_reference_file(bunny, tmp_path, key, digest_index):
    from bartels_amd import io

    path = tmp_path / f"{key}.txt"
    io.save_market(_bunny_csc(bunny, key), path)
    assert hashlib.sha256(path.read_bytes()).hexdigest() == str(bunny["sha256"][digest_index])
    back = io.load_market(path)
    assert np.array_equal(back.indptr, bunny["indptr"]) and np.array_equal(back.indices, bunny["indices"])
    assert np.array_equal(back.data.view(np.uint64), bunny[key].view(np.uint64))  # explicit zeros and -0 kept


def test_load_market_semantics(tmp_path):
    """duplicates summed in file order, rows sorted per column, symmetric files mirrored"""
    from bartels_amd import io

    p = tmp_path / "a.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n3 2 5\n3 1 1.5\n1 1 2\n3 1 -0.25\n"
                 "2 2 0\n1 2 4\n")
    A = io.load_market(p)
    assert A.shape == (3, 2)
    assert A.indptr.tolist() == [0, 2, 4] and A.indices.tolist() == [0, 2, 0, 1]
    assert A.data.tolist() == [2.0, 1.25, 4.0, 0.0]
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 2 2\n1 1 3\n2 1 -1\n")
    S = io.load_market(p).toarray()
    assert np.array_equal(S, np.array([[3.0, -1.0], [-1.0, 0.0]]))


def test_read_medit_matches_fixture(bunny):
    from bartels_amd import io

    path = os.path.join(REF_DATA, "meshes_mesh", "coarser_bunny.mesh")
    if not os.path.exists(path):
        pytest.skip("reference data not mounted")
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == str(bunny["sha256"][0])
    V, E = io.read_mesh(path)
    assert np.array_equal(V, bunny["V"]) and np.array_equal(E, bunny["E"])


def test_read_medit_sections(tmp_path):
    """the reference's two_tets.mesh layout (empty Triangles / Edges sections, indented keywords)"""
    from bartels_amd import io

    p = tmp_path / "t.mesh"
    p.write_text("MeshVersionFormatted 1\nDimension 3\nVertices\n5\n0 0 0 0\n0.1 0.1 0 0\n0.1 0 -0.1 0\n"
                 "0 0.1 -0.1 0\n0.2 0.2 -0.2 0\nTriangles\n0\n Tetrahedra\n2\n1 2 3 4 0\n2 3 4 5 0\nEdges\n0\n End\n")
    V, E = io.read_mesh(p)
    assert V.shape == (5, 3) and V[4].tolist() == [0.2, 0.2, -0.2]
    assert E.tolist() == [[0, 1, 2, 3], [1, 2, 3, 4]]


def test_read_tetgen_rules(tmp_path):
    """sim::loadTet's rules (src/read_tetgen.cpp:32-99): header counts, '#' comments and short lines
    skipped, the leading index dropped, indices kept as written, extra nodes per element ignored"""
    from bartels_amd import io

    node = tmp_path / "m.node"
    ele = tmp_path / "m.ele"
    node.write_text("4  3  0  0\n# a comment\n   0    0  0  0\n\n   1    1.5  0  0\n 2 0 2.5 0\n  3 0 0 3.5\n")
    ele.write_text("2  10  0\n# tets\n    0       0     1     2     3   9 9 9 9 9 9\n"
                   "    1       1     2     3     0   8 8 8 8 8 8\n")
    V, E = io.read_tetgen(node, ele)
    assert V.tolist() == [[0, 0, 0], [1.5, 0, 0], [0, 2.5, 0], [0, 0, 3.5]]
    assert E.tolist() == [[0, 1, 2, 3], [1, 2, 3, 0]]


def test_read_tetgen_reference_meshes():
    """the reference's TetGen meshes: counts from the headers, every index in range, and the cube's
    six tets fill its volume"""
    from bartels_amd import io
    from bartels_amd import meshgen

    cube = os.path.join(REF_DATA, "meshes_tetgen", "Cube")
    if not os.path.exists(cube):
        pytest.skip("reference data not mounted")
    V, E = io.read_tetgen(os.path.join(cube, "cube.node"), os.path.join(cube, "cube.ele"))
    assert V.shape == (8, 3) and E.shape == (6, 4)
    assert abs(np.abs(meshgen.signed_volumes(V, E)).sum() - 0.1 ** 3) < 1e-15
    beam = os.path.join(REF_DATA, "meshes_tetgen", "Beam")
    V, E = io.read_tetgen(os.path.join(beam, "Beam.node"), os.path.join(beam, "Beam.ele"))
    head_v = int(open(os.path.join(beam, "Beam.node")).readline().split()[0])
    head_e = int(open(os.path.join(beam, "Beam.ele")).readline().split()[0])
    assert V.shape == (head_v, 3) and E.shape == (head_e, 4)
    assert E.min() >= 0 and E.max() < V.shape[0]
    assert np.all(np.abs(meshgen.signed_volumes(V, E)) > 0)


def test_missing_files_are_errors(tmp_path):
    from bartels_amd import _lib, io

    with pytest.raises(_lib.BartelsError):
        io.read_mesh(tmp_path / "none.mesh")
    with pytest.raises(_lib.BartelsError):
        io.read_tetgen(tmp_path / "a.node", tmp_path / "a.ele")
    with pytest.raises(_lib.BartelsError):
        io.load_market(tmp_path / "none.mtx")

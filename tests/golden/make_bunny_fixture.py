"""Writes tests/golden/bunny.npz from the reference's own test data (data only, no code):

  data/meshes_mesh/coarser_bunny.mesh          -> V (500 x 3), E (1518 x 4, 0-based)
  data/matrices/M_coarse_bunny.txt             -> mass matrix CSC (density 0.1), used by
                                                 tests/test_assemble.cpp:25-46
  data/matrices/H_neohookean_coarser_bunny.txt -> Hessian values at q = diag(2,4,3) X, YM 6e5,
                                                 nu 0.4 (tests/test_constitutive_model.cpp:73-103)

The fixtures are stored as CSC arrays (the pattern is shared).  SHA-256 digests of the source
files are kept for provenance.  Run in a container where the reference is mounted.
"""
import hashlib
import os
import sys

import numpy as np
import scipy.io

REF = os.environ.get("BARTELS_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bunny.npz")


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    from bartels_amd.meshgen import read_mesh

    mesh = os.path.join(REF, "data", "meshes_mesh", "coarser_bunny.mesh")
    fm = os.path.join(REF, "data", "matrices", "M_coarse_bunny.txt")
    fh = os.path.join(REF, "data", "matrices", "H_neohookean_coarser_bunny.txt")
    V, E = read_mesh(mesh)
    M = scipy.io.mmread(fm).tocsc()
    H = scipy.io.mmread(fh).tocsc()
    M.sort_indices()
    H.sort_indices()
    assert np.array_equal(M.indptr, H.indptr) and np.array_equal(M.indices, H.indices)
    digests = [hashlib.sha256(open(p, "rb").read()).hexdigest() for p in (mesh, fm, fh)]
    np.savez_compressed(OUT, V=V, E=E, indptr=M.indptr.astype(np.int32), indices=M.indices.astype(np.int32),
                        M=M.data, H=H.data, sha256=np.array(digests))
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()

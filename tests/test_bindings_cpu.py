"""Host-side parts of the pysolidfem bindings that need no GPU: the
<solidfem> scene loader (PySolidFEM/Module.cpp:43-87) on the reference's own
fixture, and the boundary mesh (FemBody::BuildBoundaryMesh, FemBody.cpp:244-359)
against an independent face count, on a mesh above 65,535 nodes (where the
reference's uint16 indices would truncate)."""
import os
import sys

import numpy as np
import pytest

from solidfem_amd.mesh import BeamMesh, load_tet1

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "solidfem_amd"))
GOLD = os.path.join(ROOT, "tests", "golden")
pysolidfem = pytest.importorskip("pysolidfem")


def test_load_from_xml_reads_the_reference_fixture():
    nodes, tets, fixed = pysolidfem.load_from_xml(os.path.join(GOLD, "cantilever.xml"))
    X, T = load_tet1(os.path.join(GOLD, "box2.tet"))
    assert nodes.shape == X.shape and tets.shape == T.shape
    assert np.array_equal(nodes, X) and np.array_equal(tets, T)
    assert fixed == list(range(9))


def boundary_faces_numpy(tets, X):
    """every face of exactly one tetrahedron, with the corner opposite it"""
    faces = {}
    for e, t in enumerate(tets):
        for f in range(4):
            tri = tuple(sorted(np.delete(t, f)))
            faces.setdefault(tri, []).append((e, t[f]))
    return {k: v[0] for k, v in faces.items() if len(v) == 1}


@pytest.mark.parametrize("grid", [(4, 2, 2), (100, 25, 25)])
def test_boundary_mesh_matches_face_count_and_keeps_32_bit_ids(grid):
    m = BeamMesh(*grid)
    verts, faces, bnodes = pysolidfem.build_boundary_mesh(m.tets, m.X)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    bnodes = np.asarray(bnodes)
    if grid[0] == 4:
        want = boundary_faces_numpy(m.tets, m.X)
    else:
        # the large mesh: the boundary of a box grid is its surface nodes
        want = None
    # vertex buffer: boundary nodes in increasing order, at their (float) positions
    assert np.all(np.diff(bnodes) > 0)
    assert np.array_equal(verts, m.X[bnodes].astype(np.float32).astype(np.float64))
    tri_nodes = bnodes[faces]
    if want is not None:
        assert {tuple(sorted(t)) for t in tri_nodes} == set(want)
    # outward winding: the normal points away from the box's inside
    a, b, c = (m.X[tri_nodes[:, k]] for k in range(3))
    n = np.cross(b - a, c - a)
    centre = m.X.mean(axis=0)
    assert np.all(np.einsum("ij,ij->i", n, (a + b + c) / 3 - centre) > 0)
    # every surface node of the box grid is a boundary vertex, none other
    lo, hi = m.X.min(axis=0), m.X.max(axis=0)
    on_surface = np.any(np.isclose(m.X, lo) | np.isclose(m.X, hi), axis=1)
    assert np.array_equal(np.sort(bnodes), np.flatnonzero(on_surface))
    if grid[0] == 100:
        assert m.nn > 65535 and bnodes.max() > 65535
        nx, ny, nz = grid
        assert faces.shape[0] == 4 * (nx * ny + ny * nz + nx * nz)   # 2 triangles per cell face

"""Scene formats on the CPU: the reference's own cantilever.xml / box2.tet
fixtures parse to the mesh, Dirichlet nodes and FemConfig values the
reference's IO::LoadFromXmlFile gives them (Source/FemIO.cpp:583-847), and the
VTK writer round-trips."""
import os

import numpy as np
import pytest

from solidfem_amd import io

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_cantilever_xml_scene():
    sc = io.load_xml(os.path.join(GOLD, "cantilever.xml"))
    assert sc["nodes"].shape == (81, 3) and sc["tets"].shape == (192, 4)
    assert list(sc["fixed"]) == list(range(9))
    c = sc["config"]
    assert c["material"] == 3 and c["young"] == 66000 and c["poisson"] == 0.45 and c["density"] == 1070
    assert c["simtype"] == 1 and c["force_step"] == 1.0 and c["gravity"] == -9.8
    assert c["maxiter"] == 100 and c["tol"] == 0.1
    # box2.tet swaps y and z on read: the file's third column becomes y
    assert np.isclose(sc["nodes"][1, 1], 0.05) and sc["nodes"][1, 2] == 0.0
    # the fixed nodes are the x = 0 face of the box
    assert np.all(sc["nodes"][:9, 0] == 0.0)


def test_xml_variants(tmp_path):
    tet = tmp_path / "m.tet"
    tet.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nt 0 1 2 3\n")
    xml = tmp_path / "s.xml"
    xml.write_text('<solidfem><model path="m.tet" type="tet1" scale="2" swap="xy"><dirichlet>0, 1</dirichlet></model>'
                   '<material name="neo-hookean" young="1e5"/>'
                   '<simulation type="explicit" steps="4" method="nonlinear" gravity="-3">'
                   '<pressure p="250">0 1 2</pressure></simulation></solidfem>')
    sc = io.load_xml(str(xml))
    # tet1 swaps y/z, then the model's swap="xy" exchanges x and y; scale 2
    assert np.allclose(sc["nodes"][1], [0, 2, 0]) and np.allclose(sc["nodes"][3], [2, 0, 0])
    assert list(sc["fixed"]) == [0, 1] and list(sc["surf_tris"]) == [0, 1, 2]
    c = sc["config"]
    assert c["simtype"] == 2 and c["substeps"] == 4 and "force_step" not in c
    assert c["pressure"] == 250 and c["gravity"] == -3 and c["young"] == 1e5
    bad = tmp_path / "b.xml"
    bad.write_text('<solidfem><model path="m.vol" type="vol"/></solidfem>')
    with pytest.raises(ValueError):
        io.load_xml(str(bad))


def test_vtk_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    X0 = rng.standard_normal((7, 3))
    x = X0 + 0.1 * rng.standard_normal((7, 3))
    tets = np.array([[0, 1, 2, 3], [3, 4, 5, 6]])
    P = rng.standard_normal((2, 3, 3))
    io.export_vtk(str(tmp_path / "a.vtk"), X0, tets, x, P)
    r = io.read_vtk(str(tmp_path / "a.vtk"))
    assert np.array_equal(r["points"], x) and np.array_equal(r["cells"], tets)
    assert np.allclose(r["displacement"], x - X0, rtol=1e-8) and np.allclose(r["stress"], P, rtol=1e-8)


def test_prepare_orients_every_tet_positively():
    sc = io.load_xml(os.path.join(GOLD, "cantilever.xml"))
    X = sc["nodes"]

    def vols(T):
        return np.linalg.det(np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=2)) / 6

    assert np.all(vols(sc["tets"]) < 0)          # the y/z swap on read mirrors box2.tet
    T = io.prepare(X, sc["tets"])
    assert np.all(vols(T) > 0)
    assert np.array_equal(T[:, 2:], sc["tets"][:, 2:])
    assert np.allclose(np.abs(vols(T)), np.abs(vols(sc["tets"])))

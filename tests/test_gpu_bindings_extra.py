"""The rest of the reference's pysolidfem surface on the GPU (PySolidFEM/
Module.cpp:43-111): boundary and visual meshes and OBJ export after steps,
get_hessian, the force-parameter gradients, cables from a <solidfem> scene,
and ComputeEnergy -- each against the FP64 oracle or an independent
computation.  Meshes above 65,535 nodes keep their boundary (the reference's
uint16 truncation, FemBody.cpp:263,310-313, is not reproduced)."""
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd.io import prepare
from solidfem_amd.mesh import BeamMesh, create_cable, load_tet1

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "solidfem_amd"))
GOLD = os.path.join(ROOT, "tests", "golden")
pytestmark = pytest.mark.gpu


def _fem(m, cfg):
    import pysolidfem
    fixed = np.nonzero(m.fixed_mask())[0].astype(np.int32)
    base = dict(young=1e6, poisson=0.45, density=1070.0, gravity=-9.8, zup=True)
    base.update(cfg)
    return pysolidfem.NonlinearFEM(m.tets, m.X, fixed, base)


def shuffled_free(m):
    """free nodes in the reference's shuffled order (fixed first, then free,
    each in original order, FemPhysicsMatrixFree.cpp:229-263)"""
    return np.flatnonzero(m.fixed_mask() == 0)


def read_obj(path):
    v, f = [], []
    for line in open(path):
        t = line.split()
        if t and t[0] == "v":
            v.append([float(x) for x in t[1:4]])
        elif t and t[0] == "f":
            f.append([int(x.split("//")[0]) - 1 for x in t[1:4]])
    return np.array(v), np.array(f)


def test_boundary_mesh_above_65535_nodes_and_obj_after_a_step(tmp_path):
    print("case boundary mesh, 68,276-node beam; 10x3x3 step + OBJ", flush=True)
    m = BeamMesh(100, 25, 25)
    fem = _fem(m, dict(material="linear", solver="newton_ls", rel_tol=1e-8, tol=0.0))
    verts, faces, bnodes = fem.get_boundary_mesh()
    assert m.nn > 65535 and max(bnodes) > 65535 and len(faces) == 3 * 4 * (100 * 25 + 25 * 25 + 100 * 25)
    assert np.array_equal(verts, m.X[bnodes].astype(np.float32).astype(np.float64))
    del fem
    m = BeamMesh(10, 3, 3)
    fem = _fem(m, dict(material="neo-hookean", solver="newton_ls", rel_tol=1e-8, tol=0.0))
    fem.step()
    verts, faces, bnodes = fem.get_boundary_mesh()
    x = fem.get_nodes()
    assert np.array_equal(verts, x[bnodes].astype(np.float32).astype(np.float64))
    assert np.abs(verts - m.X[bnodes]).max() > 1e-3       # the mesh moved with the step
    path = str(tmp_path / "b.obj")
    fem.save_to_obj(path)
    v, f = read_obj(path)
    assert np.abs(v - verts).max() < 1e-6 and np.array_equal(f.reshape(-1), np.asarray(faces))


def test_visual_mesh_and_cable_from_a_scene(tmp_path):
    """a <solidfem> scene with a visual OBJ and a <cable>: the visual mesh is
    mapped to the boundary and follows it; the cable's forces enter the
    residual (checked through the oracle's FP64 residual at the solution)"""
    import pysolidfem
    print("case scene with visual mesh and cable", flush=True)
    shutil.copy(os.path.join(GOLD, "box2.tet"), tmp_path / "box2.tet")
    X, T = load_tet1(os.path.join(GOLD, "box2.tet"))
    vb, fb, nb = pysolidfem.build_boundary_mesh(T, X)
    centre = X.mean(axis=0)
    vis = centre + 0.95 * (np.asarray(vb) - centre)          # a slightly smaller copy of the surface
    with open(tmp_path / "vis.obj", "w") as out:
        for p in vis:
            out.write(f"v {p[0]:.9f} {p[1]:.9f} {p[2]:.9f}\n")
        for t in np.asarray(fb).reshape(-1, 3):
            out.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
    (tmp_path / "scene.xml").write_text(
        '<solidfem>\n <model path="box2.tet" type="tet1" visual="vis.obj">\n'
        '  <dirichlet>0 1 2 3 4 5 6 7 8</dirichlet>\n </model>\n'
        ' <material name="neo-hookean" young="66000" poisson="0.45" density="1070"/>\n'
        ' <simulation type="quasi-static" steps="2" method="nonlinear" gravity="-9.8"/>\n'
        ' <solver method="newton" iterations="100" tolerance="1e-6"/>\n'
        ' <cable divs="6" len="0.3" offset="0.05,0.05,0.05" axis="x" stiffness="2000"/>\n</solidfem>\n')
    fem = pysolidfem.NonlinearFEM(str(tmp_path / "scene.xml"))
    assert fem.num_cables() == 1
    v0, f0, _ = fem.get_visual_mesh()
    assert np.abs(v0 - vis).max() < 1e-6
    fem.set_cable_actuation(0, 0.8)
    fem.step()
    fem.step()
    assert fem.last_stats()["converged"] == 1
    x = fem.get_nodes()
    # the visual vertices moved with the surface they are tied to
    v1, _, _ = fem.get_visual_mesh()
    moved = np.linalg.norm(v1 - v0, axis=1)
    assert moved.max() > 1e-3
    # the FP64 oracle's residual with the same cable at the device solution,
    # on the mesh as the bindings prepare it (FemBody::Prepare: the y/z swap of
    # .tet files mirrors every tetrahedron, whose first two indices are swapped)
    T = prepare(X, T)
    c = create_cable(X, T, divs=6, length=0.3, offset=(0.05, 0.05, 0.05), stiffness=2000.0)
    oc = O.OracleCable(T, c["elem"], c["bary"], c["rest_length"], 2000.0, actuation=0.8)
    fixed = np.zeros(len(X), np.uint8)
    fixed[:9] = 1
    p = O.OracleProblem(X, T, fixed, O.NEO_HOOKEAN, 66000.0, 0.45, 1070.0, gravity=(0, -9.8, 0))
    assert np.abs(fem.get_cable_positions(0) - oc.positions(x)).max() < 1e-12
    r = p.gradients(x, 1.0) + oc.forces(x, 1.0) * (1 - fixed[:, None])
    # the scene's absolute tolerance, as the device's FP64 residual met it
    assert np.linalg.norm(r) < 1e-6 * 1.0001, np.linalg.norm(r)
    assert np.abs(oc.forces(x, 1.0)).max() > 1.0      # the cable pulls


def test_get_hessian_matches_oracle_columns():
    print("case get_hessian 6x2x2", flush=True)
    m = BeamMesh(6, 2, 2)
    fem = _fem(m, dict(material="stvk", solver="newton_ls", rel_tol=1e-8, tol=0.0))
    fem.step()
    x = fem.get_nodes()
    K = fem.get_hessian()
    free = shuffled_free(m)
    n = 3 * len(free)
    assert K.shape == (n, n)
    p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), O.STVK, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    Ko = np.zeros((n, n))
    for j in range(n):
        d = np.zeros_like(m.X)
        d[free[j // 3], j % 3] = 1.0
        Ko[:, j] = p.matvec(x, d)[free].reshape(-1)
    err = np.linalg.norm(K - Ko) / np.linalg.norm(Ko)
    print(f"  hessian rel err {err:.2e}", flush=True)
    assert err < 2e-6
    assert np.abs(K - K.T).max() < 1e-5 * np.abs(K).max()


def test_force_param_grads_through_bindings():
    print("case compute_force_param_grads", flush=True)
    m = BeamMesh(10, 3, 3)
    fem = _fem(m, dict(material="neo-hookean", solver="newton_ls", rel_tol=1e-8, tol=0.0))
    fem.step()
    x = fem.get_nodes()
    fem.compute_force_param_grads()
    free = shuffled_free(m)
    p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    fmu, flam = p.force_param_grads(x)
    frho = p.nodal_volumes()[:, None] * np.array([0.0, 0.0, -9.8])
    for got, want in ((fem.get_force_mu_grad(), fmu), (fem.get_force_lambda_grad(), flam), (fem.get_force_rho_grad(), frho)):
        want = want[free].reshape(-1)
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_compute_energy_matches_reference_formula():
    """ComputeEnergy with gravity, the pressure term and a cable, against the
    oracle's term-by-term restatement (pinned to the compiled reference's
    ComputeEnergy in tests/test_oracle_cables_energy.py)"""
    print("case compute_energy", flush=True)
    m = BeamMesh(10, 3, 3)
    for zup in (True, False):
        fem = _fem(m, dict(material="neo-hookean", solver="newton_ls", rel_tol=1e-8, tol=0.0, zup=zup,
                           simtype=1, force_step=0.5))
        tris = m.tip_triangles()
        fem.set_pressure(tris, 1e4)
        c = create_cable(m.X, m.tets, divs=10, length=1.0, offset=(0.0, 0.125, 0.125))
        fem.add_cable(c["elem"], c["bary"], c["rest_length"], c["stiffness"])
        fem.set_cable_actuation(0, 0.9)
        fem.step()
        x = fem.get_nodes()
        g = (0.0, 0.0, -9.8) if zup else (0.0, -9.8, 0.0)
        p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=g)
        oc = O.OracleCable(m.tets, c["elem"], c["bary"], c["rest_length"], c["stiffness"], actuation=0.9)
        want = O.reference_total_energy(p, x, 0.5, g, tris=tris, pressure=1e4, cables=[oc])
        got = fem.compute_energy()
        print(f"  zup {zup}: energy {got:.10e} oracle {want:.10e}", flush=True)
        assert abs(got - want) <= 1e-11 * abs(want)
        assert abs(fem.compute_elastic_energy() - p.energy(x)) <= 1e-11 * p.energy(x)

"""Explicit dynamics, element stresses and the scene formats on the GPU.

- sfem_step in ST_EXPLICIT against O.explicit_steps (the SubStep restatement,
  itself pinned to the compiled reference in test_oracle.py);
- sfem_element_stress against the oracle's per-element P(F);
- the reference's cantilever.xml scene (box2.tet) solved on the device and
  checked against the oracle's FP64 solve, then exported to VTK."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi, io
from solidfem_amd.mesh import BeamMesh, lame, uniform01

pytestmark = pytest.mark.gpu
MATS = [capi.LINEAR, capi.STVK, capi.COROTATIONAL, capi.NEO_HOOKEAN]
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.mark.parametrize("mat", MATS)
def test_explicit_substeps_match_oracle(mat):
    print(f"case explicit mat={mat}", flush=True)
    m = BeamMesh(6, 2, 3, 0.6, 0.2, 0.3)
    fixed = m.fixed_mask()
    g = (0.0, 0.0, -9.8)
    s = capi.Solver(m.X, m.tets, fixed, mat, 2e5, 0.3, 900.0, gravity=g)
    p = O.OracleProblem(m.X, m.tets, fixed, mat, 2e5, 0.3, 900.0, gravity=g)
    prm = capi.solve_params(sim_type=capi.ST_EXPLICIT, time_step=6e-3, substeps=3)
    for _ in range(4):
        st = s.step(prm)
        assert st["force_evals"] == 3 * m.ne
    x, v = O.explicit_steps(p, m.X, np.zeros_like(m.X), 2e-3, 12, g)
    assert rel(s.positions() - m.X, x - m.X) < 1e-11
    assert rel(s.velocities(), v) < 1e-11
    assert np.all(s.positions()[fixed == 1] == m.X[fixed == 1])
    s.reset()
    assert np.all(s.velocities() == 0)


@pytest.mark.parametrize("mat", [capi.LINEAR, capi.STVK, capi.NEO_HOOKEAN])
def test_implicit_step_matches_dense_oracle(mat):
    """Converged backward-Euler steps (ST_IMPLICIT) against O.ImplicitStep's
    dense FP64 Newton (pinned to the compiled reference's residual and
    operator in test_oracle.py); then the dynamic terms stay in sfem_gradients
    / sfem_matvec, as mTimeStep stays set in the reference."""
    print(f"case implicit mat={mat}", flush=True)
    m = BeamMesh(4, 2, 2)
    fixed = m.fixed_mask()
    g = (0.0, 0.0, -9.8)
    h = 4e-3
    s = capi.Solver(m.X, m.tets, fixed, mat, 1e5, 0.3, 1000.0, gravity=g)
    p = O.OracleProblem(m.X, m.tets, fixed, mat, 1e5, 0.3, 1000.0, gravity=g)
    prm = capi.solve_params(tol=1e-11, sim_type=capi.ST_IMPLICIT, time_step=2 * h, substeps=1, inner=100000)
    x, v = m.X.copy(), np.zeros_like(m.X)
    for k in range(3):
        st = s.step(prm)
        assert st["converged"] == 1
        imp = O.ImplicitStep(p, x, v, 2 * h)
        xn, _ = imp.solve(x)
        v = (xn - x) / (2 * h)
        x = xn
        assert rel(s.positions() - m.X, x - m.X) < 1e-8, k
        assert rel(s.velocities(), v) < 1e-7, k
    # state after the last step: x0 = its start, v = its velocity update
    x0 = x - 2 * h * v
    imp = O.ImplicitStep(p, x0, v, 2 * h)
    xs = s.positions()
    assert rel(s.gradients(), imp.residual(xs)) < 1e-10
    d = np.random.default_rng(mat).standard_normal(m.X.shape)
    assert rel(s.matvec(d), imp.matvec(xs, d)) < 2e-6
    s.reset()
    assert rel(s.matvec(d), p.matvec(m.X, d)) < 2e-6      # static operator again


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not shipped")
def test_implicit_reference_mode_matches_compiled_reference():
    """REFERENCE mode (SolveNewtonCG: fixed Newton steps, CG to 1e-3) inside
    implicit Step(dt), against the compiled reference doing the same."""
    m = BeamMesh(6, 2, 3, 0.6, 0.2, 0.3)
    fixed = m.fixed_mask()
    ref = O.RefProblem(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 2e5, 0.3, 900.0, -9.8, zup=True,
                       simtype=O.RefProblem.ST_IMPLICIT, outer=10, inner=100)
    s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 2e5, 0.3, 900.0, gravity=(0, 0, -9.8))
    prm = capi.solve_params(mode=capi.SOLVE_REFERENCE, outer=10, inner=100, sim_type=capi.ST_IMPLICIT, time_step=5e-3,
                            substeps=1)
    for _ in range(3):
        ref.step_dt(5e-3)
        s.step(prm)
    xr = ref.positions()
    err = rel(s.positions() - m.X, xr - m.X)
    print(f"implicit reference mode vs compiled reference: {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("mat", MATS)
@pytest.mark.parametrize("per_elem", [False, True])
def test_element_stress_matches_oracle(mat, per_elem):
    print(f"case stress mat={mat} per_elem={per_elem}", flush=True)
    m = BeamMesh(8, 3, 3)
    fixed = m.fixed_mask()
    rng = np.random.default_rng(mat)
    if per_elem:
        E = 10 ** (5 + 2 * uniform01(4, 1, m.ne))
        nu = 0.30 + 0.18 * uniform01(4, 2, m.ne)
        mu, lam = lame(E, nu)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, young_e=E, poisson_e=nu)
    else:
        mu, lam = lame(np.full(m.ne, 1e6), np.full(m.ne, 0.45))
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0)
    x = m.X + 0.05 * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    s.set_positions(x)
    P = s.element_stress()
    T = m.tets
    for e in range(0, m.ne, 7):
        Dm = (m.X[T[e, 1:]] - m.X[T[e, 0]]).T
        Ds = (x[T[e, 1:]] - x[T[e, 0]]).T
        F = Ds @ np.linalg.inv(Dm)
        Po = O.stress(mat, mu[e], lam[e], F)
        assert np.abs(P[e] - Po).max() <= 1e-10 * np.abs(Po).max() + 1e-9, e


def test_cantilever_xml_scene_solves_and_exports(tmp_path):
    sc = io.load_xml(os.path.join(GOLD, "cantilever.xml"))
    c = sc["config"]
    X, T = sc["nodes"], io.prepare(sc["nodes"], sc["tets"])   # as FemBody::LoadFromXml does
    fixed = np.zeros(len(X), dtype=np.uint8)
    fixed[sc["fixed"]] = 1
    g = (0.0, c["gravity"], 0.0)          # FemConfig::mZUpAxis defaults to false: gravity along y
    s = capi.Solver(X, T, fixed, c["material"], c["young"], c["poisson"], c["density"], gravity=g)
    st = s.step(capi.solve_params(tol=1e-10, sim_type=c["simtype"], force_step=c["force_step"], inner=100000))
    assert st["converged"] == 1
    p = O.OracleProblem(X, T, fixed, c["material"], c["young"], c["poisson"], c["density"], gravity=g)
    xo, sto = p.newton_solve(X, 1.0, tol=1e-10)
    assert sto["converged"]
    x = s.positions()
    print(f"cantilever.xml: tip sag {(x - X)[:, 1].min():.4e} m, err vs oracle {rel(x - X, xo - X):.2e}")
    assert rel(x - X, xo - X) < 1e-6
    out = tmp_path / "cantilever.vtk"
    io.export_vtk(str(out), X, T, x, s.element_stress())
    r = io.read_vtk(str(out))
    assert np.array_equal(r["points"], x) and np.array_equal(r["cells"], T)
    assert np.allclose(r["displacement"], x - X, rtol=1e-7, atol=1e-12)

"""The pysolidfem bindings (NonlinearFEM) and the C++ FemPhysicsMatrixFree
class they wrap, on the GPU: reference-mode Newton-CG against the compiled
reference's golden vector, converged Newton-LS against the FP64 oracle, and
the quasi-static driver `simulate_static`."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd.mesh import BeamMesh

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "solidfem_amd"))

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fem(m, cfg):
    import pysolidfem
    fixed = np.nonzero(m.fixed_mask())[0].astype(np.int32)
    base = dict(young=1e6, poisson=0.45, density=1070.0, gravity=-9.8, zup=True)
    base.update(cfg)
    return pysolidfem.NonlinearFEM(m.tets, m.X, fixed, base)


def test_reference_newton_cg_through_bindings():
    g = np.load(os.path.join(GOLD, "ref_cfg1_mat3.npz"))
    m = BeamMesh(10, 3, 3)
    fem = _fem(m, dict(material="neo-hookean", solver="newton_cg", maxiter=10, inner_iters=1000))
    fem.set_pressure(m.tip_triangles(), 1e4)
    fem.step()
    st = fem.last_stats()
    assert st["newton_iters"] == 10
    u, ur = fem.get_nodes() - m.X, g["newton_x"] - m.X
    err = np.linalg.norm(u - ur) / np.linalg.norm(ur)
    print("bindings reference-mode err", err)
    assert err < 3e-7, err
    assert np.array_equal(fem.get_tets(), m.tets)


@pytest.mark.parametrize("material", ["linear", "stvk", "neo-hookean"])
def test_converged_newton_ls_through_bindings(material):
    m = BeamMesh(10, 3, 3)
    fem = _fem(m, dict(material=material, solver="newton_ls", rel_tol=1e-8, tol=0.0, maxiter=50))
    fem.step()
    st = fem.last_stats()
    assert st["converged"] == 1
    mat = {"linear": O.LINEAR, "stvk": O.STVK, "neo-hookean": O.NEO_HOOKEAN}[material]
    p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), mat, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    xr, _ = p.newton_solve(None, 1.0, tol=1e-10, cg_rtol=1e-8)
    err = np.linalg.norm(fem.get_nodes() - xr) / np.linalg.norm(xr - m.X)
    assert err < 1e-6


def test_simulate_static_quasi_static_and_lame():
    m = BeamMesh(12, 3, 3)
    fem = _fem(m, dict(material="neo-hookean", solver="newton_ls", rel_tol=1e-8, tol=0.0, simtype=1,
                       force_step=0.25))
    fem.simulate_static()
    assert abs(fem.get_force_fraction() - 1.0) < 1e-15
    assert fem.last_stats()["converged"] == 1
    assert not fem.check_for_inversion()
    mu, lam = fem.get_shear_modulus(), fem.get_lame_lambda()
    assert abs(mu - 0.5e6 / 1.45) < 1e-6 and abs(lam - 1e6 * 0.45 / 1.45 / 0.1) < 1e-4
    fem.set_lame_params(2 * mu, 2 * lam)
    assert abs(fem.get_shear_modulus() - 2 * mu) < 1e-6


def test_xml_scene_through_bindings(tmp_path):
    """NonlinearFEM(path) on the reference's own cantilever.xml / box2.tet:
    Neo-Hookean + NST_NEWTON_LS as the reference's path constructor sets them,
    gravity along y (mZUpAxis false), |r| <= 0.1 N (the scene's tolerance);
    the result against the oracle's FP64 solve, then save_to_vtk."""
    import pysolidfem
    from solidfem_amd import io
    fem = pysolidfem.NonlinearFEM(os.path.join(GOLD, "cantilever.xml"))
    fem.simulate_static()
    st = fem.last_stats()
    assert st["converged"] == 1 and st["res"] <= 0.1
    sc = io.load_xml(os.path.join(GOLD, "cantilever.xml"))
    X, T = sc["nodes"], io.prepare(sc["nodes"], sc["tets"])
    fixed = np.zeros(len(X), dtype=np.uint8)
    fixed[sc["fixed"]] = 1
    c = sc["config"]
    p = O.OracleProblem(X, T, fixed, O.NEO_HOOKEAN, c["young"], c["poisson"], c["density"], gravity=(0, -9.8, 0))
    x = fem.get_nodes()
    # the device stopped at |r| <= 0.1 N: the FP64 oracle's residual at the
    # device state says the same, and the state is near the oracle's converged
    # one (stopping at 0.1 N leaves ~2e-3 relative, as the oracle itself shows)
    r_oracle = np.linalg.norm(p.gradients(x, 1.0))
    xo, sto = p.newton_solve(X, 1.0, tol=1e-10)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo - X)
    print(f"cantilever.xml via bindings: sag {(x - X)[:, 1].min():.4e} m, err {err:.2e}, |r| device "
          f"{st['res']:.3e} oracle {r_oracle:.3e}")
    assert abs(r_oracle - st["res"]) <= 1e-6 * max(st["res"], 1e-3) and r_oracle <= 0.1
    assert err < 2e-2
    out = tmp_path / "c.vtk"
    fem.save_to_vtk(str(out))
    r = io.read_vtk(str(out))
    assert np.allclose(r["points"], x, rtol=0, atol=1e-15) and np.array_equal(r["cells"], T)
    assert np.allclose(r["displacement"], x - X, rtol=1e-6, atol=1e-12)
    assert r["stress"].shape == (len(T), 3, 3) and np.isfinite(r["stress"]).all()


def test_explicit_through_bindings():
    """config simtype 2 (ST_EXPLICIT) with substeps: step(dt) runs the device
    SubStep; against the oracle's restatement."""
    m = BeamMesh(6, 2, 3, 0.6, 0.2, 0.3)
    fem = _fem(m, dict(simtype=2, substeps=4, material="stvk", young=2e5, poisson=0.3, density=900.0))
    for _ in range(3):
        fem.step(8e-3)
    p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), O.STVK, 2e5, 0.3, 900.0, gravity=(0, 0, -9.8))
    x, _ = O.explicit_steps(p, m.X, np.zeros_like(m.X), 2e-3, 12, (0.0, 0.0, -9.8))
    assert np.linalg.norm(fem.get_nodes() - x) / np.linalg.norm(x - m.X) < 1e-11


def test_multigrid_through_bindings():
    """precond "multigrid" with the levels added through add_multigrid_level:
    the converged Newton-LS solve against the FP64 oracle, fewer CG iterations
    than block Jacobi, and the levels kept across set_lame_params"""
    from solidfem_amd.mesh import beam_hierarchy
    m = BeamMesh(16, 8, 8)
    cfg = dict(material="neo-hookean", solver="newton_ls", rel_tol=1e-8, tol=0.0, maxiter=50)
    fem = _fem(m, dict(cfg, precond="multigrid"))
    levels = beam_hierarchy(m)          # 8x4x4, 675 dofs: solved directly
    for lv in levels:
        fem.add_multigrid_level(*lv)
    assert fem.num_multigrid_levels() == 1 + len(levels) == 2
    fem.step()
    st = fem.last_stats()
    fj = _fem(m, dict(cfg, precond="jacobi"))
    fj.step()
    stj = fj.last_stats()
    p = O.OracleProblem(m.X, m.tets, m.fixed_mask(), O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    xr, _ = p.newton_solve(None, 1.0, tol=1e-10, cg_rtol=1e-8)
    err = np.linalg.norm(fem.get_nodes() - xr) / np.linalg.norm(xr - m.X)
    print(f"bindings multigrid: err {err:.2e} cg {st['cg_iters']} jacobi cg {stj['cg_iters']}")
    assert st["converged"] == 1 and st["fallbacks"] == 0 and err < 1e-6
    assert st["cg_iters"] < 0.5 * stj["cg_iters"]
    fem.set_lame_params(fem.get_shear_modulus(), fem.get_lame_lambda())
    assert fem.num_multigrid_levels() == 2
    fem.clear_multigrid()
    assert fem.num_multigrid_levels() == 1


def test_multigrid_binding_errors():
    from solidfem_amd.mesh import beam_hierarchy
    m = BeamMesh(16, 4, 4)
    with pytest.raises(RuntimeError):
        _fem(m, dict(precond="amg"))
    fem = _fem(m, dict(solver="newton_ls", rel_tol=1e-8, tol=0.0, precond="multigrid"))
    X0, tets, fx, inj, ptr, col, val = beam_hierarchy(m)[0]
    with pytest.raises(RuntimeError):
        fem.add_multigrid_level(X0, tets, fx, inj, ptr[:-1], col, val)   # rows do not cover the fine mesh
    with pytest.raises(RuntimeError):
        fem.add_multigrid_level(X0, tets, fx, inj, ptr, col, val, young_e=np.ones(len(tets)))
    assert fem.num_multigrid_levels() == 1
    with pytest.raises(RuntimeError):
        fem.step()                                                       # no coarse levels

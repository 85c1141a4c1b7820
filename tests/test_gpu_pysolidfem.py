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
    assert err < 1e-4
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

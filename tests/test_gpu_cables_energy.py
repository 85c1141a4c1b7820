"""Cable springs, energies and force-parameter gradients on the GPU through
the C-ABI, against the FP64 oracle (itself pinned to the compiled reference in
tests/test_oracle_cables_energy.py).  FP64 device kernels: 1e-11 relative."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, create_cable, lame, uniform01

pytestmark = pytest.mark.gpu
MATS_ALL = [capi.LINEAR, capi.STVK, capi.COROTATIONAL, capi.NEO_HOOKEAN, capi.COROTATIONAL_POLAR]


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(mat, grid=(10, 3, 3), per_elem=False, seed=0):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    kw = dict(gravity=(0.0, 0.0, -9.8))
    rng = np.random.default_rng(seed)
    fdead = rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    if per_elem:
        E = 10 ** (5 + 2 * uniform01(seed, 1, m.ne))
        nu = 0.30 + 0.18 * uniform01(seed, 2, m.ne)
        mu, lam = lame(E, nu)
        p = O.OracleProblem(m.X, m.tets, fixed, mat, mu=mu, lam=lam, density=1070.0, fext=fdead, **kw)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, young_e=E, poisson_e=nu, **kw)
    else:
        p = O.OracleProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, fext=fdead, **kw)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, **kw)
    s.set_external_forces(fdead)
    x = m.X + 0.03 * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    x[:, 2] -= 0.05 * m.X[:, 0] ** 2
    return m, fixed, p, s, x


def add_cable(m, s, actuation=0.85):
    c = create_cable(m.X, m.tets, divs=10, length=1.0, offset=(0.0, 0.125, 0.125))
    cid = s.add_cable(c["elem"], c["bary"], c["rest_length"], c["stiffness"])
    s.set_cable_actuation(cid, actuation)
    return O.OracleCable(m.tets, c["elem"], c["bary"], c["rest_length"], c["stiffness"], actuation=actuation)


@pytest.mark.parametrize("mat", MATS_ALL)
@pytest.mark.parametrize("per_elem", [False, True])
def test_energy_matches_oracle(mat, per_elem):
    print(f"case energy mat={mat} per_elem={per_elem}", flush=True)
    m, fixed, p, s, x = setup(mat, (12, 4, 4), per_elem, seed=mat + 3)
    s.set_positions(x)
    s.set_load_fraction(0.7)
    e_el, e_sp, w = s.energy()
    print(f"  elastic {e_el:.6e} (oracle {p.energy(x):.6e}) work {w:.6e}", flush=True)
    assert abs(e_el - p.energy(x)) <= 1e-11 * abs(p.energy(x))
    assert e_sp == 0.0
    assert abs(w - p.dead_work(x, 0.7)) <= 1e-11 * abs(p.dead_work(x, 0.7))
    # deterministic
    assert s.energy() == (e_el, e_sp, w)


@pytest.mark.parametrize("mat", MATS_ALL)
def test_force_param_grads_match_oracle(mat):
    print(f"case force param grads mat={mat}", flush=True)
    m, fixed, p, s, x = setup(mat, seed=mat + 5)
    s.set_positions(x)
    fmu, flam, frho = s.force_param_grads()
    omu, olam = p.force_param_grads(x)
    assert relmax(fmu, omu) < 1e-11 and relmax(flam, olam) < 1e-11
    want = p.nodal_volumes()[:, None] * np.array([0.0, 0.0, -9.8]) * (1 - fixed[:, None])
    assert relmax(frho, want) < 1e-12


@pytest.mark.parametrize("mat", [capi.LINEAR, capi.NEO_HOOKEAN])
def test_residual_and_energy_with_cable_match_oracle(mat):
    print(f"case cable residual mat={mat}", flush=True)
    m, fixed, p, s, x = setup(mat, seed=11)
    oc = add_cable(m, s, 0.85)
    s.set_positions(x)
    s.set_load_fraction(0.6)
    r = s.gradients()
    ro = p.gradients(x, 0.6) + oc.forces(x, 0.6) * (1 - fixed[:, None])
    assert np.abs(oc.forces(x, 0.6)).max() > 1e-3 * np.abs(ro).max()
    assert relmax(r, ro) < 1e-11, relmax(r, ro)
    assert np.all(r[fixed == 1] == 0)
    e_el, e_sp, _ = s.energy()
    assert abs(e_sp - oc.energy(x)) <= 1e-11 * abs(oc.energy(x))
    assert np.abs(s.cable_positions(0, 11) - oc.positions(x)).max() < 1e-14
    # actuation 0 switches the cable off; an index out of range is ignored
    s.set_cable_actuation(0, 0.0)
    s.set_cable_actuation(7, 0.5)
    assert relmax(s.gradients(), p.gradients(x, 0.6)) < 1e-11
    assert s.energy()[1] == 0.0 and s.num_cables() == 1
    s.clear_cables()
    assert s.num_cables() == 0


def test_quasi_static_solve_with_cable_meets_fp64_oracle_residual():
    """a contracting cable bends the beam: Newton-PCG converges to 1e-8 in the
    oracle's FP64 residual (elastic + gravity + cable); the matrix-free tangent
    has no cable stiffness, as in the reference, so Newton is inexact"""
    print("case cable quasi-static solve", flush=True)
    m, fixed, p, s, _ = setup(capi.NEO_HOOKEAN, (16, 4, 4), seed=2)
    s.set_external_forces(None)
    p = O.OracleProblem(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    c = create_cable(m.X, m.tets, divs=16, length=1.0, offset=(0.0, 0.125, 0.05))
    s.add_cable(c["elem"], c["bary"], c["rest_length"], 5e4)
    s.set_cable_actuation(0, 0.9)
    oc = O.OracleCable(m.tets, c["elem"], c["bary"], c["rest_length"], 5e4, actuation=0.9)
    for k in range(4):
        st = s.step(tol=1e-8, precond=capi.PC_JACOBI, sim_type=capi.ST_QUASI_STATIC, force_step=0.25, outer=200)
        print(f"  step {k}: newton {st['newton_iters']} cg {st['cg_iters']}", flush=True)
        assert st["converged"] == 1
    x = s.positions()
    r = p.gradients(x, 1.0) + oc.forces(x, 1.0) * (1 - fixed[:, None])
    ext = p.external(x, 1.0) + oc.forces(x, 1.0) * (1 - fixed[:, None])
    res = np.linalg.norm(r) / np.linalg.norm(ext)
    lift = (x - m.X)[:, 2].max()
    print(f"  oracle rel residual {res:.2e}, tip motion {lift:.3e} m", flush=True)
    assert res < 1e-8


def test_explicit_substeps_with_cable_match_oracle():
    print("case cable explicit", flush=True)
    m, fixed, p, s, _ = setup(capi.STVK, seed=4)
    s.set_external_forces(None)
    p = O.OracleProblem(m.X, m.tets, fixed, capi.STVK, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    oc = add_cable(m, s, 0.7)
    h, nsub = 1e-4, 20
    s.step(capi.solve_params(sim_type=capi.ST_EXPLICIT, time_step=h * nsub, substeps=nsub))
    x, v = m.X.copy(), np.zeros_like(m.X)
    free = fixed == 0
    for _ in range(nsub):
        f = p.internal_forces(x) + oc.forces(x, 1.0, v)
        v[free] += (h / p.mass[free])[:, None] * f[free] + h * np.array([0.0, 0.0, -9.8])
        x[free] += h * v[free]
    err = np.linalg.norm(s.positions() - x) / np.linalg.norm(x - m.X)
    print(f"  explicit with cable rel err {err:.2e}", flush=True)
    assert err < 1e-10


def test_free_cable_nodes_and_bad_elements_are_refused():
    m = BeamMesh(4, 2, 2)
    s = capi.Solver(m.X, m.tets, m.fixed_mask(), capi.LINEAR, 1e6, 0.3, 1000.0)
    with pytest.raises(capi.SfemError) as e:
        s.add_cable([0, -3, 2], np.full((3, 3), 0.25), 0.1)
    assert "free cable" in str(e.value)
    with pytest.raises(capi.SfemError):
        s.add_cable([0, m.ne], np.full((2, 3), 0.25), 0.1)
    assert s.num_cables() == 0

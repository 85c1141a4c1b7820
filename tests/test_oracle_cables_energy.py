"""Pins the oracle's cable springs, energies and force-parameter gradients:
against the compiled reference (oracle/_ref: FemPhysicsBase::
ComputeSpringForces through ComputeGradients, ComputeEnergy,
ComputeElasticEnergy, ComputeSpringEnergy, GetForceParamGrads) where it is
built, and against mathematics everywhere (forces as minus the energy
gradient by central differences, the linearity of the stress in mu and
lambda, the parameter derivatives by central differences)."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd.mesh import BeamMesh, create_cable

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built")
MATS = [O.LINEAR, O.STVK, O.COROTATIONAL, O.NEO_HOOKEAN]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def deformed(m, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    fixed = m.fixed_mask()
    x = m.X + amp * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    x[:, 2] -= 0.05 * m.X[:, 0] ** 2      # a bent beam: the cable along the axis stretches
    return x


def cable_on(m, actuation=0.8):
    c = create_cable(m.X, m.tets, divs=10, length=1.0, offset=(0.0, 0.125, 0.125))
    assert not c["free"].any()
    return c, O.OracleCable(m.tets, c["elem"], c["bary"], c["rest_length"], c["stiffness"], actuation=actuation)


def test_create_cable_points_and_weights():
    """the cable's points are where CreateCable puts them (offset + i len/divs
    dir) and its weights reproduce them on the rest mesh"""
    m = BeamMesh(10, 3, 3)
    c, oc = cable_on(m)
    pts = oc.positions(m.X)
    want = np.column_stack([np.linspace(0.0, 1.0, 11), np.full(11, 0.125), np.full(11, 0.125)])
    assert np.abs(pts - want).max() < 1e-14
    assert np.all(oc.w4 > -1e-12)


@pytest.mark.parametrize("act", [0.8, 1.0 - 0.005, 1.2])
def test_cable_forces_are_minus_energy_gradient(act):
    """stretched (act 0.8), inside the smoothed band |err| <= rest/100 (0.995),
    and slack (1.2, no force) springs"""
    m = BeamMesh(10, 3, 3)
    x = deformed(m, 1, amp=0.0) if act != 0.995 else m.X.copy()
    if act == 1.0 - 0.005:
        x = m.X.copy()
    _, oc = cable_on(m, act)
    f = oc.forces(x, 1.0)
    g = np.zeros_like(x)
    h = 1e-7
    for i in np.unique(oc.node4):
        for q in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, q] += h
            xm[i, q] -= h
            g[i, q] = -(oc.energy(xp) - oc.energy(xm)) / (2 * h)
    if act == 1.2:
        assert np.all(f == 0) and oc.energy(x) == 0
    else:
        assert np.abs(f).max() > 0
        assert rel(f, g) < 1e-6, rel(f, g)
    # a spring's forces cancel: the cable applies no net force
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * max(np.abs(f).max(), 1e-300)


def test_cable_forces_scale_with_load_fraction_and_vanish_unactuated():
    m = BeamMesh(10, 3, 3)
    x = deformed(m, 2)
    _, oc = cable_on(m)
    assert np.allclose(oc.forces(x, 0.3), 0.3 * oc.forces(x, 1.0), rtol=1e-14, atol=0)
    oc.actuation = 0.0
    assert np.all(oc.forces(x, 1.0) == 0) and oc.energy(x) == 0


@needs_ref
@pytest.mark.parametrize("mat", [O.LINEAR, O.NEO_HOOKEAN])
def test_cable_residual_matches_compiled_reference(mat):
    """ComputeGradients with a cable (FemPhysicsMatrixFree.cpp:536-560): the
    oracle's residual plus its cable forces equals the compiled reference's"""
    m = BeamMesh(10, 3, 3)
    fixed = m.fixed_mask()
    c, oc = cable_on(m, 0.85)
    ref = O.RefProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, -9.8, zup=True)
    ref.add_cable(c["elem"], c["bary"], c["rest_length"], c["stiffness"], actuation=0.85)
    x = deformed(m, 3)
    ref.set_positions(x)
    ref.set_fraction(0.6)
    p = O.OracleProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    r = p.gradients(x, 0.6) + oc.forces(x, 0.6) * (1 - fixed[:, None])
    rr = ref.gradients()
    assert np.abs(oc.forces(x, 0.6)).max() > 1e-3 * np.abs(rr).max()
    assert rel(r, rr) < 1e-12, rel(r, rr)


@needs_ref
@pytest.mark.parametrize("mat", MATS)
@pytest.mark.parametrize("zup", [True, False])
def test_energies_match_compiled_reference(mat, zup):
    """ComputeElasticEnergy, ComputeSpringEnergy and ComputeEnergy (with its
    gravity, pressure and spring terms) of the compiled reference"""
    m = BeamMesh(10, 3, 3)
    fixed = m.fixed_mask()
    tris = m.tip_triangles()
    c, oc = cable_on(m, 0.9)
    ref = O.RefProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, -9.8, zup=zup, tris=tris, pressure=1e4)
    ref.add_cable(c["elem"], c["bary"], c["rest_length"], c["stiffness"], actuation=0.9)
    x = deformed(m, 4)
    ref.set_positions(x)
    ref.set_fraction(0.7)
    g = (0.0, 0.0, -9.8) if zup else (0.0, -9.8, 0.0)
    p = O.OracleProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, gravity=g)
    assert abs(p.energy(x) - ref.elastic_energy()) <= 1e-12 * abs(ref.elastic_energy())
    assert abs(oc.energy(x) - ref.spring_energy()) <= 1e-12 * abs(ref.spring_energy())
    e = O.reference_total_energy(p, x, 0.7, g, tris=tris, pressure=1e4, cables=[oc])
    assert abs(e - ref.energy()) <= 1e-12 * abs(ref.energy()), (e, ref.energy())


@pytest.mark.parametrize("mat", MATS + [O.COROTATIONAL_POLAR])
def test_force_param_grads_by_finite_differences(mat):
    """d f_int / d mu and d f_int / d lambda by central differences of the
    internal force in each modulus"""
    m = BeamMesh(6, 2, 2)
    fixed = m.fixed_mask()
    x = deformed(m, 5, amp=0.05)
    p = O.OracleProblem(m.X, m.tets, fixed, mat, mu=3e5, lam=7e5)
    fmu, flam = p.force_param_grads(x)
    for k, fa in ((0, fmu), (1, flam)):
        h = 1.0
        pp = O.OracleProblem(m.X, m.tets, fixed, mat, mu=3e5 + (h if k == 0 else 0), lam=7e5 + (h if k == 1 else 0))
        pm = O.OracleProblem(m.X, m.tets, fixed, mat, mu=3e5 - (h if k == 0 else 0), lam=7e5 - (h if k == 1 else 0))
        fd = (pp.internal_forces(x) - pm.internal_forces(x)) / (2 * h)
        assert rel(fa, fd) < 1e-7, (k, rel(fa, fd))
    # linearity: f = mu fmu + lam flam
    assert rel(3e5 * fmu + 7e5 * flam, p.internal_forces(x)) < 1e-12


@needs_ref
def test_force_param_grads_match_compiled_reference():
    """GetForceParamGrads (FemPhysicsBase.cpp:162-187) for the material whose
    ComputeElementStress fills Pmu and Plambda (Neo-Hookean), and the density
    gradient (nodal volume times gravity)"""
    m = BeamMesh(10, 3, 3)
    fixed = m.fixed_mask()
    ref = O.RefProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, -9.8, zup=True)
    x = deformed(m, 6)
    ref.set_positions(x)
    gmu, glam, grho = ref.force_param_grads()
    p = O.OracleProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8))
    fmu, flam = p.force_param_grads(x)
    free = fixed == 0
    assert rel(fmu[free], gmu[free]) < 1e-12 and rel(flam[free], glam[free]) < 1e-12
    want_rho = p.nodal_volumes()[:, None] * np.array([0.0, 0.0, -9.8])
    assert rel(want_rho[free], grho[free]) < 1e-12

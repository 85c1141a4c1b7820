"""GPU tests of the Neo-Hookean tangent cache (sfem_set_tangent_cache), through
the C-ABI.

With the cache on, each Newton step builds the tangent's state per element
(E = Dm^-1 F^-1 and V (mu - lam ln J)) once and every Hessian-vector product of
its CG solve is made from it, in a different but algebraically equal
arrangement of the reference's tangent (FemPhysicsMatrixFree.cpp:464-470); the
multigrid PCG does not use it (the switch must leave its solves bitwise
unchanged).
Tolerances: the cached product agrees with the FP64 oracle as the rebuilding
product does (2e-6 relative in the 2-norm, 2e-5 with per-element moduli) and
with the rebuilding product to FP32 rounding (5e-7, measured ~1e-7); converged
solves agree with the oracle to 1e-6 and with each other to 1e-7."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, lame, traction_nodal_forces, uniform01

pytestmark = pytest.mark.gpu


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(grid, per_elem=False, seed=3, mat=capi.NEO_HOOKEAN):
    m = BeamMesh(*grid, 1.0, 0.25, 0.25)
    fixed = m.fixed_mask()
    tris = m.tip_triangles()
    kw = dict(gravity=(0.0, 0.0, -9.8))
    if per_elem:
        E = 10 ** (5 + 2 * uniform01(seed, 1, m.ne))
        nu = 0.30 + 0.18 * uniform01(seed, 2, m.ne)
        mu, lam = lame(E, nu)
        p = O.OracleProblem(m.X, m.tets, fixed, mat, mu=mu, lam=lam, density=1070.0, tris=tris, pressure=1e4, **kw)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, young_e=E, poisson_e=nu, **kw)
    else:
        p = O.OracleProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, tris=tris, pressure=1e4, **kw)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, **kw)
    s.set_pressure_surface(tris, 1e4)
    return m, fixed, p, s


@pytest.mark.parametrize("grid,per_elem,amp", [((10, 3, 3), False, 0.05), ((24, 6, 8), True, 0.05),
                                               ((32, 8, 8), False, 0.08)])
def test_cached_product_matches_oracle(grid, per_elem, amp):
    print(f"case grid={grid} per_elem={per_elem} amp={amp}", flush=True)
    m, fixed, p, s = setup(grid, per_elem)
    rng = np.random.default_rng(7)
    x = m.X + amp * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    x[:, 2] -= 0.1 * m.X[:, 0] ** 2          # a bent state: J and F^-1 far from the identity's
    d = rng.standard_normal(m.X.shape)
    s.set_positions(x)
    s.set_tangent_cache(0)
    q0 = s.matvec(d)
    s.set_tangent_cache(1)                  # solves only: matvec unchanged, bit for bit
    q1 = s.matvec(d)
    s.set_tangent_cache(2)
    q2 = s.matvec(d)
    q2b = s.matvec(d)
    qo = p.matvec(x, d)
    tol = 2e-5 if per_elem else 2e-6
    e0, e2, e02 = rel2(q0, qo), rel2(q2, qo), rel2(q2, q0)
    print(f"  rebuilt {e0:.2e}, cached {e2:.2e} vs oracle; cached vs rebuilt {e02:.2e}", flush=True)
    assert np.array_equal(q1, q0)
    assert np.array_equal(q2b, q2)
    assert np.all(q2[fixed == 1] == 0)
    assert e2 < tol and e02 < (5e-6 if per_elem else 5e-7), (e2, e02)


def test_cache_applies_to_neo_hookean_only():
    """other materials ignore the switch: their products are bitwise unchanged"""
    for mat in (capi.LINEAR, capi.STVK, capi.COROTATIONAL, capi.COROTATIONAL_POLAR):
        m, fixed, p, s = setup((10, 3, 3), mat=mat)
        rng = np.random.default_rng(mat)
        s.set_positions(m.X + 0.05 * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None]))
        d = rng.standard_normal(m.X.shape)
        s.set_tangent_cache(0)
        q0 = s.matvec(d)
        s.set_tangent_cache(2)
        assert np.array_equal(s.matvec(d), q0), mat


def beam(grid, multigrid):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    tris = m.tip_triangles()
    f = traction_nodal_forces(m.X, tris, lambda q: np.tile([0.0, 0.0, -1e4], (len(q), 1)))
    s.set_external_forces(f)
    if multigrid:
        s.set_multigrid(beam_hierarchy(m))
    return m, fixed, f, s


@pytest.mark.parametrize("precond", [capi.PC_JACOBI, capi.PC_NONE, capi.PC_MULTIGRID])
def test_cached_newton_solve_matches_rebuilt_and_oracle(precond):
    """two quasi-static load steps of the 32x8x8 beam with and without the
    cache: both converge, agree to 1e-7 and with the FP64 oracle's Newton
    solution to 1e-6; the multigrid solves are bitwise the same"""
    print(f"case 32x8x8 two load steps precond={precond}", flush=True)
    out = []
    for mode in (0, 1):
        m, fixed, f, s = beam((32, 8, 8), precond == capi.PC_MULTIGRID)
        s.set_tangent_cache(mode)
        stats = []
        for k in range(2):
            st = s.step(tol=1e-8, precond=precond, sim_type=capi.ST_QUASI_STATIC, force_step=0.5)
            stats.append((st["converged"], st["newton_iters"], st["cg_iters"], st["fallbacks"]))
        out.append((stats, s.positions() - m.X))
    print(f"  rebuilt {out[0][0]}, cached {out[1][0]}", flush=True)
    assert all(t[0] == 1 and t[3] == 0 for t in out[0][0] + out[1][0])
    du = rel2(out[1][1], out[0][1])
    p = O.OracleProblem(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8), fext=f)
    xo = m.X.copy()
    for k in range(2):
        xo, _ = p.newton_solve(xo, 0.5 * (k + 1), tol=1e-10, cg_rtol=1e-8)
    eo = rel2(out[1][1], xo - m.X)
    print(f"  cached vs rebuilt {du:.2e}, cached vs oracle {eo:.2e}", flush=True)
    assert du < 1e-7 and eo < 1e-6, (du, eo)
    if precond == capi.PC_MULTIGRID:
        assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])


def test_mode_checked():
    m, fixed, p, s = setup((10, 3, 3))
    with pytest.raises(capi.SfemError):
        capi._check(capi.lib().sfem_set_tangent_cache(s.h, 3))

"""GPU tests of the multigrid V-cycle's folded launch schedule
(sfem_mg_set_schedule), through the C-ABI.

The folded schedule makes each level's Chebyshev steps inside the element
kernels' row staging, the residual inside the restriction, the prolongation
inside the next operator apply, the finest level's first step inside the CG
update and the r.z reduction inside the last step.  Every node value is the
same FP32 expression as in the one-kernel-per-step schedule, so the two give
bitwise the same V-cycle, linear solve and Newton solve; these tests compare
them bit for bit on hierarchies with one and two smoothed levels, a directly
solved or a smoothed coarsest level, per-element moduli and each material."""
import numpy as np
import pytest

from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, traction_nodal_forces, uniform01

pytestmark = pytest.mark.gpu


def setup(grid, material=capi.NEO_HOOKEAN, per_element=False, min_cells=2):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    kw = {}
    hk = {}
    if per_element:
        E = 10 ** (5 + 2 * uniform01(7, 1, m.ne))
        nu = 0.30 + 0.18 * uniform01(7, 2, m.ne)
        kw = dict(young_e=E, poisson_e=nu)
        hk = dict(young_e=E, poisson_e=nu)
    s = capi.Solver(m.X, m.tets, fixed, material, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8), **kw)
    tris = m.tip_triangles()
    f = traction_nodal_forces(m.X, tris, lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1)))
    s.set_external_forces(f)
    s.set_multigrid(beam_hierarchy(m, min_cells=min_cells, **hk))
    x = m.X.copy()
    x[:, 2] -= 0.05 * m.X[:, 0] ** 2
    s.set_positions(x)
    return m, fixed, s


CASES = [
    ("nh 32x8x8", (32, 8, 8), capi.NEO_HOOKEAN, False, 2),
    ("nh 64x16x16 two smoothed levels", (64, 16, 16), capi.NEO_HOOKEAN, False, 2),
    ("nh 64x16x16 smoothed coarsest 32x8x8", (64, 16, 16), capi.NEO_HOOKEAN, False, 8),
    ("nh per-element moduli 32x8x8", (32, 8, 8), capi.NEO_HOOKEAN, True, 2),
    ("linear 32x8x8", (32, 8, 8), capi.LINEAR, False, 2),
    ("stvk 32x8x8", (32, 8, 8), capi.STVK, False, 2),
    ("corotational 32x8x8", (32, 8, 8), capi.COROTATIONAL, False, 2),
    ("polar corotational 32x8x8", (32, 8, 8), capi.COROTATIONAL_POLAR, False, 2),
]


@pytest.mark.parametrize("name,grid,material,per_element,min_cells", CASES, ids=[c[0] for c in CASES])
def test_folded_vcycle_and_pcg_bitwise_equal(name, grid, material, per_element, min_cells):
    print(f"case {name}", flush=True)
    m, fixed, s = setup(grid, material, per_element, min_cells)
    rng = np.random.default_rng(11)
    r = rng.standard_normal(m.X.shape) * (1 - fixed)[:, None]
    # the V-cycle at the bent state; PCG at rest, where the tangent is linear
    # elasticity (at the bent state, an off-equilibrium compressed one, the
    # Neo-Hookean tangent need not be positive definite)
    xb = s.positions()
    s.mg_set_schedule(False)
    z0 = s.mg_apply(r)
    s.set_positions(m.X)
    x0, it0, st0 = s.mg_pcg(r, rtol=1e-6, maxit=200)
    s.mg_set_schedule(True)
    x1, it1, st1 = s.mg_pcg(r, rtol=1e-6, maxit=200)
    s.set_positions(xb)
    z1 = s.mg_apply(r)
    z1b = s.mg_apply(r)
    nl = len(s.mg_info())
    print(f"levels {nl} |z| {np.linalg.norm(z0):.6e} max|dz| {np.abs(z1 - z0).max():.3e} "
          f"pcg {it0}/{it1} status {st0}/{st1}", flush=True)
    assert np.all(np.isfinite(z0)) and np.linalg.norm(z0) > 0
    assert np.array_equal(z1, z0)
    assert np.array_equal(z1b, z1)
    assert (it1, st1) == (it0, st0) and st0 == 1
    assert np.array_equal(x1, x0)


@pytest.mark.parametrize("precond", [capi.PC_MULTIGRID, capi.PC_MULTIGRID_LEARNED])
def test_folded_schedule_newton_solve_bitwise_equal(precond):
    """two quasi-static load steps of the 32x8x8 NH beam (with the load-step
    predictor and the judge) give the same iterations and bitwise the same
    displacements under both schedules"""
    print(f"case 32x8x8 two load steps precond={precond}", flush=True)
    out = []
    for fused in (False, True):
        m = BeamMesh(32, 8, 8)
        fixed = m.fixed_mask()
        s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
        tris = m.tip_triangles()
        s.set_external_forces(traction_nodal_forces(m.X, tris, lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1))))
        s.set_multigrid(beam_hierarchy(m))
        s.mg_set_schedule(fused)
        if precond == capi.PC_MULTIGRID_LEARNED:
            from solidfem_amd import learned
            net = learned.load(smoother=True)
            net.attach(s)
        stats = []
        for k in range(2):
            st = s.step(tol=1e-8, precond=precond, sim_type=capi.ST_QUASI_STATIC, force_step=0.5)
            stats.append((st["converged"], st["newton_iters"], st["cg_iters"], st["fallbacks"]))
        out.append((stats, s.positions()))
    print(f"unfolded {out[0][0]} folded {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0] and all(t[0] == 1 for t in out[0][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_pcg_outcomes_iteration_limit_and_convergence():
    """the multigrid PCG's host loop reads each iteration's outcome from the
    slot its direction kernel writes (flag, iteration count) and may hold the
    next iteration back when convergence is predicted: an iteration limit of
    k returns k iterations and status 0 (the same iterate as the first k
    iterations of a longer solve, bitwise), and a converged solve reports the
    count at which it converged, whatever the limit beyond it"""
    m, fixed, s = setup((32, 8, 8))
    s.set_positions(m.X)
    rng = np.random.default_rng(5)
    r = rng.standard_normal(m.X.shape) * (1 - fixed)[:, None]
    xc, itc, stc = s.mg_pcg(r, rtol=1e-6, maxit=200)
    print(f"converged: {itc} iterations, status {stc}", flush=True)
    assert stc == 1 and 2 < itc < 200
    x2, it2, st2 = s.mg_pcg(r, rtol=1e-6, maxit=itc + 1)
    assert (it2, st2) == (itc, 1) and np.array_equal(x2, xc)
    for k in (1, 2, itc - 1):
        xk, itk, stk = s.mg_pcg(r, rtol=1e-6, maxit=k)
        print(f"maxit {k}: {itk} iterations, status {stk}", flush=True)
        assert (itk, stk) == (k, 0)
        assert np.linalg.norm(xk - xc) > 0
    xa, _, _ = s.mg_pcg(r, rtol=1e-6, maxit=itc - 1)
    assert np.array_equal(xa, xk)


def test_schedule_argument_checked():
    m, fixed, s = setup((16, 4, 4))
    with pytest.raises(capi.SfemError):
        capi._check(capi.lib().sfem_mg_set_schedule(s.h, 2))


def test_attached_network_outlives_its_temporary():
    """the context keeps a raw pointer to an attached network: the solver
    object holds a reference to it, so attaching a temporary leaves the
    network alive until it is detached"""
    import gc
    import weakref

    from solidfem_amd import learned
    m, fixed, s = setup((16, 4, 4))
    learned.load(smoother=True).attach(s)
    ref = weakref.ref(s._learned)
    gc.collect()
    assert ref() is not None and ref().h
    capi.LearnedPreconditioner.detach(s)
    gc.collect()
    assert ref() is None

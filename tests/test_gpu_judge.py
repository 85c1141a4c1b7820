"""The fallback and the judge of the converged Newton-PCG, forced from the host.

A learned or multigrid PCG that breaks down (p.Kp <= 0, r.z <= 0), stagnates,
hits its iteration cap or meets a non-finite value is redone with block-Jacobi
PCG for that Newton step (`fallbacks`); a learned or multigrid solve whose
final FP64 residual misses the target is rerun from its start state with
Jacobi-PCG (`reruns`, the judge).  Each case here installs preconditioner
blocks that cannot work (NaN, negative definite) or caps the preconditioned
attempt's Newton steps, and checks that the solve still converges to the FP64
oracle's solution within 1e-6 (the reference's own Newton guards against
non-finite and growing residuals are at Include/NewtonSolver.h:156-172)."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, traction_nodal_forces

pytestmark = pytest.mark.gpu


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def beam(grid=(16, 4, 4), mg=False):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    f = traction_nodal_forces(m.X, m.tip_triangles(), lambda p: np.tile([0.0, 0.0, -1e3], (len(p), 1)))
    s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    s.set_external_forces(f)
    if mg:
        s.set_multigrid(beam_hierarchy(m))
    p = O.OracleProblem(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8), fext=f)
    return m, fixed, s, p


def check_against_oracle(m, s, p):
    xo, ost = p.newton_solve(None, 1.0, tol=1e-10, cg_rtol=1e-8)
    x = s.positions()
    err = rel2(x - m.X, xo - m.X)
    res = np.linalg.norm(p.gradients(x, 1.0)) / np.linalg.norm(p.external(x, 1.0))
    return err, res


@pytest.mark.parametrize("blocks", ["nan", "negative"])
def test_learned_blocks_that_cannot_work_fall_back_to_jacobi(blocks):
    print(f"case learned blocks {blocks}: every linear solve falls back", flush=True)
    m, fixed, s, p = beam()
    minv = np.zeros((m.nn, 9), dtype=np.float32)
    if blocks == "nan":
        minv[:] = np.nan
    else:
        minv[:, [0, 4, 8]] = -1.0 / 1e5     # negative definite: r.z < 0 at once
    s.set_precond_blocks(minv)
    s.set_load_fraction(1.0)
    st = s.solve(tol=1e-8, precond=capi.PC_LEARNED)
    err, res = check_against_oracle(m, s, p)
    print(f"  newton {st['newton_iters']} cg {st['cg_iters']} fallbacks {st['fallbacks']} reruns {st['reruns']} "
          f"err {err:.2e} oracle-res {res:.2e}", flush=True)
    assert st["converged"] == 1
    assert st["fallbacks"] == st["newton_iters"] > 0 and st["reruns"] == 0
    assert res < 1e-8 and err < 1e-6


def test_multigrid_with_nan_smoother_blocks_falls_back_to_jacobi():
    """NaN learned blocks in the finest smoother: the V-cycle's set-up is not
    finite, so each Newton step's linear solve runs block-Jacobi PCG"""
    print("case multigrid_learned with NaN blocks", flush=True)
    m, fixed, s, p = beam(mg=True)
    s.set_precond_blocks(np.full((m.nn, 9), np.nan, dtype=np.float32))
    s.set_load_fraction(1.0)
    st = s.solve(tol=1e-8, precond=capi.PC_MULTIGRID_LEARNED)
    err, res = check_against_oracle(m, s, p)
    print(f"  newton {st['newton_iters']} cg {st['cg_iters']} fallbacks {st['fallbacks']} err {err:.2e}", flush=True)
    assert st["converged"] == 1 and st["fallbacks"] == st["newton_iters"] > 0
    assert res < 1e-8 and err < 1e-6


@pytest.mark.parametrize("precond", [capi.PC_MULTIGRID, capi.PC_LEARNED])
def test_judge_reruns_a_solve_that_misses_the_fp64_target(precond):
    """the preconditioned attempt gets one Newton step (pc_newton_cap = 1), so
    its FP64 residual misses 1e-8: the judge reruns the solve from its start
    state with Jacobi-PCG, which converges"""
    print(f"case judge rerun precond={precond}", flush=True)
    m, fixed, s, p = beam(mg=precond == capi.PC_MULTIGRID)
    if precond == capi.PC_LEARNED:
        net = capi.LearnedPreconditioner(seed=3)   # untrained: block Jacobi
        net.attach(s)
    s.set_load_fraction(1.0)
    st = s.solve(capi.solve_params(tol=1e-8, precond=precond, pc_newton_cap=1))
    err, res = check_against_oracle(m, s, p)
    print(f"  newton {st['newton_iters']} cg {st['cg_iters']} fallbacks {st['fallbacks']} reruns {st['reruns']} "
          f"err {err:.2e} oracle-res {res:.2e}", flush=True)
    assert st["converged"] == 1 and st["reruns"] == 1 and st["fallbacks"] >= 1
    assert res < 1e-8 and err < 1e-6
    # without the cap the preconditioned solve converges on its own
    s.reset()
    s.set_load_fraction(1.0)
    st2 = s.solve(tol=1e-8, precond=precond)
    assert st2["converged"] == 1 and st2["reruns"] == 0 and st2["fallbacks"] == 0


def test_multigrid_pcg_stops_on_breakdown_or_stagnation():
    """an indefinite operator (a compressed Neo-Hookean state past its
    stability limit) makes one V-cycle-PCG solve break down or stagnate within
    the stagnation window instead of running to the iteration cap"""
    print("case mg_pcg on an indefinite state", flush=True)
    m, fixed, s, p = beam(mg=True)
    # axial compression of 40%: the tangent of the beam is indefinite (buckling)
    x = m.X.copy()
    x[:, 0] *= 0.6
    s.set_positions(x)
    b = np.random.default_rng(2).standard_normal(m.X.shape) * (1 - fixed[:, None])
    _, it, status = s.mg_pcg(b, rtol=1e-12, maxit=400)
    print(f"  iters {it} status {status}", flush=True)
    assert status in (3, 5) or (status == 1 and it < 60)
    assert it <= 80

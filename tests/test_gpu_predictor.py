"""GPU tests of the quasi-static load-step predictors (sfem_step, CONVERGED
mode): 0 = the last state (what the reference's Step does), 1 = the secant
extrapolation of the last two converged states, 2 = the quadratic through the
last three, with the secant one, then the last state, as fallbacks when a
predicted start is rejected.  A predictor only changes where Newton starts:
every load step must still end at the FP64 oracle's converged state."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, traction_nodal_forces

pytestmark = pytest.mark.gpu


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run(grid, nu, jitter, precond, predictor, steps):
    m = BeamMesh(*grid, jitter=jitter, seed=5)
    fixed = m.fixed_mask()
    s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, nu, 1070.0, gravity=(0.0, 0.0, -9.8))
    f = traction_nodal_forces(m.X, m.tip_triangles(), lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1)))
    s.set_external_forces(f)
    if precond in (capi.PC_MULTIGRID, capi.PC_MULTIGRID_LEARNED):
        s.set_multigrid(beam_hierarchy(m))
    hist, states = [], []
    for k in range(steps):
        st = s.step(tol=1e-8, precond=precond, sim_type=capi.ST_QUASI_STATIC, force_step=1.0 / steps,
                    predictor=predictor)
        hist.append((st["converged"], st["newton_iters"], st["cg_iters"], st["rejects"], st["fallbacks"]))
        states.append(s.positions())
    return m, fixed, f, hist, states


CASES = [
    ("nh 32x8x8 jacobi", (32, 8, 8), 0.45, 0.0, capi.PC_JACOBI, 5),
    ("nh 32x8x8 multigrid", (32, 8, 8), 0.45, 0.0, capi.PC_MULTIGRID, 5),
    ("nh nu=0.495 jittered 32x8x8 multigrid", (32, 8, 8), 0.495, 0.1, capi.PC_MULTIGRID, 4),
]


@pytest.mark.parametrize("name,grid,nu,jitter,precond,steps", CASES, ids=[c[0] for c in CASES])
def test_predictors_reach_the_oracle_state(name, grid, nu, jitter, precond, steps):
    print(f"case {name}", flush=True)
    runs = {}
    for pred in (0, 1, 2):
        m, fixed, f, hist, states = run(grid, nu, jitter, precond, pred, steps)
        runs[pred] = (hist, states)
        print(f"  predictor {pred}: (converged, newton, cg, rejects, fallbacks) per load step {hist}", flush=True)
        assert all(h[0] == 1 for h in hist)
    p = O.OracleProblem(m.X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, nu, 1070.0, gravity=(0.0, 0.0, -9.8), fext=f)
    xo = m.X.copy()
    for k in range(steps):
        xo, _ = p.newton_solve(xo, (k + 1) / steps, tol=1e-10, cg_rtol=1e-8)
        for pred in (0, 1, 2):
            x = runs[pred][1][k]
            err = rel2(x - m.X, xo - m.X)
            assert err < 1e-6, (pred, k, err)
    errs = {pred: rel2(runs[pred][1][-1] - m.X, xo - m.X) for pred in (0, 1, 2)}
    print(f"  final error vs the FP64 oracle: {errs}", flush=True)
    newton = {pred: sum(h[1] for h in runs[pred][0]) for pred in (0, 1, 2)}
    print(f"  newton steps {newton}", flush=True)
    # the predictors never cost more than the reference's start from the last
    # state by more than a step per load step (rejected predictions included)
    assert newton[2] <= newton[0] + steps and newton[1] <= newton[0] + steps


def test_quadratic_predictor_is_deterministic():
    out = [run((16, 4, 4), 0.45, 0.0, capi.PC_MULTIGRID, 2, 4) for _ in range(2)]
    assert out[0][3] == out[1][3]
    assert all(np.array_equal(a, b) for a, b in zip(out[0][4], out[1][4]))

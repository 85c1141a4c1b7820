"""The five workloads (solidfem_amd/configs.py) on the GPU against the FP64
oracle: reduced grids with the full physics of each workload, and the full
grids of workloads 2, 3 (polar corotational), 4 and 5 with the FP64 residual
of the device solution checked by the oracle and the displacement error
measured (FP64 Newton polish) or estimated (the FP64 residual's Newton
correction).  SFEM_FULL=1 adds the remaining full-size combinations
(minutes)."""
import json
import os

import numpy as np
import pytest

from solidfem_amd import capi, configs
from tests.cfg_util import oracle_problem

pytestmark = pytest.mark.gpu
FULL = os.environ.get("SFEM_FULL") == "1"


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_solution(d, tol=1e-10, max_newton=60):
    p = oracle_problem(d, nthreads=16)
    x = d["mesh"].X.copy()
    steps = d["load_steps"]
    for k in range(steps):
        x, st = p.newton_solve(x, (k + 1) / steps, tol=tol, cg_rtol=0.1, max_newton=max_newton, max_cg=200000)
    return x, st, p


@pytest.mark.parametrize("cfg,grid", [(1, None), (2, (16, 4, 4)), (4, (16, 4, 4)), (5, (16, 4, 4))])
def test_config_converged_against_oracle(cfg, grid):
    out, d, s = configs.run(cfg, grid=grid)
    x = s.positions()
    xo, st, p = oracle_solution(d)
    err = rel2(x - d["mesh"].X, xo - d["mesh"].X)
    res = np.linalg.norm(p.gradients(x, 1.0)) / np.linalg.norm(p.external(x, 1.0))
    print(f"config {cfg} grid {grid}: newton {out['newton_iters']} cg {out['cg_iters']} "
          f"oracle-res {res:.2e} err vs FP64 oracle {err:.2e}")
    assert out["converged"] and st["converged"]
    assert res < 1e-8
    assert err < 1e-6


def test_config3_corotational_reference_semantics():
    """The reference's corotational tangent is not the Jacobian of its force
    and not symmetric, so its CG (fixed 1e-3 forcing, no convergence test on
    the Newton residual) amplifies operator rounding: the FP64 oracle itself
    moves by the spread printed below when its moduli change by +-2e-7, about
    FP32 rounding (a CG run that stops one iteration earlier or later).  The
    device and the oracle run the same fixed-step algorithm (reference mode, 4
    Newton steps per load step) on a reduced grid; the device must agree with
    the oracle to within a small multiple of that spread."""
    d = configs.build(3, (16, 4, 4))
    s = configs.solver(d)
    X = d["mesh"].X

    def oracle_run(scale):
        p = oracle_problem(d, modulus_scale=scale)
        x = X.copy()
        for k in range(4):
            x, _ = p.newton_cg_ref(x, 0.05 * (k + 1), 4, 2000)
        return x

    for k in range(4):
        s.step(capi.solve_params(mode=capi.SOLVE_REFERENCE, outer=4, inner=2000, check_every=8,
                                 sim_type=capi.ST_QUASI_STATIC, force_step=0.05))
    x = oracle_run(1.0)
    err = rel2(s.positions() - X, x - X)
    spread = max(rel2(oracle_run(1.0 + sc) - X, x - X) for sc in (2e-7, -2e-7))
    print(f"config 3 reduced, reference semantics: err {err:.2e}, oracle spread under +-2e-7 moduli {spread:.2e}")
    assert err < max(10 * spread, 1e-5)


# full-size workloads run by default: (workload, preconditioner); "3p" is
# workload 3 with the polar corotational model.  SFEM_FULL=1 adds the rest.
FULL_DEFAULT = [(2, "jacobi"), (2, "multigrid_learned"), (4, "jacobi"), (4, "multigrid_learned"), ("3p", "multigrid"),
                (5, "multigrid")]
FULL_OPT_IN = [(1, "jacobi"), (3, "jacobi"), ("3p", "jacobi"), (5, "multigrid_learned"), (5, "jacobi")]


def newton_correction_estimate(s, d, x, p):
    """||K(u)^-1 r(u)|| / ||u||: the FP64 oracle's residual r at the device
    state u, and the Newton correction it asks for, solved by the device's
    multigrid PCG to 1e-3 relative at that state.  In Newton's quadratic
    regime this is the distance to the FP64 solution on the same branch (to a
    relative O(||correction||)), the error estimate where an FP64 CPU solve is
    out of reach (the 6.3M-tet nu = 0.495 beam)."""
    from solidfem_amd.mesh import beam_hierarchy
    r = p.gradients(x, 1.0)
    if s.mg_info().size <= 1:
        s.set_multigrid(beam_hierarchy(d["mesh"], young_e=d["young_e"], poisson_e=d["poisson_e"]))
    dx, it, status = s.mg_pcg(r, rtol=1e-3, maxit=400)
    X = d["mesh"].X
    return float(np.linalg.norm(dx) / np.linalg.norm(x - X)), it, status


@pytest.mark.parametrize("cfg,pc", FULL_DEFAULT + (FULL_OPT_IN if FULL else []))
def test_full_size_config(cfg, pc):
    """Full grids.  The oracle checks the FP64 residual of the device solution;
    for 1 it solves the problem itself; for 2, 3p and 4 it continues Newton in
    FP64 from the device solution to 1e-10 (same equilibrium branch) and the
    displacement error is measured against that; for every workload the
    error is also estimated by the FP64 residual's Newton correction (the only
    measure for 5: an FP64 CPU solve at nu = 0.495 takes hours), which the
    oracle polish of 2, 3p and 4 validates.  3 uses the reference's
    corotational tangent, with which Newton stalls (no 1e-8 claim)."""
    print(f"case full-size workload {cfg} precond {pc}", flush=True)
    pcv = configs.PRECONDS[pc]
    mat = capi.COROTATIONAL_POLAR if cfg == "3p" else None
    cfg = 3 if cfg == "3p" else cfg
    stall = cfg == 3 and mat is None
    out, d, s = configs.run(cfg, pcv, max_newton=12 if stall else 60, log=lambda m: print(m, flush=True), material=mat)
    out["precond"] = pc
    x = s.positions()
    p = oracle_problem(d, nthreads=16)
    out["oracle_rel_residual"] = float(np.linalg.norm(p.gradients(x, 1.0)) / np.linalg.norm(p.external(x, 1.0)))
    X = d["mesh"].X
    if cfg == 1:
        xo, st, _ = oracle_solution(d)
        out["err_vs_fp64_oracle"] = float(rel2(x - X, xo - X))
    elif cfg in (2, 4) or (cfg == 3 and not stall):
        # one FP64 Newton step per call, so that each step reports as it ends
        xo, tot = x, {"newton": 0, "cg": 0, "converged": 0}
        for k in range(10):
            xo, st = p.newton_solve(xo, 1.0, tol=1e-10, cg_rtol=1e-2, max_newton=1, max_cg=200000)
            tot["newton"] += st["newton"]
            tot["cg"] += st["cg"]
            tot["converged"] = st["converged"]
            print(f"oracle polish step {k + 1}: cg {st['cg']} converged {st['converged']}", flush=True)
            if st["converged"]:
                break
        out["err_vs_fp64_oracle"] = float(rel2(x - X, xo - X))
        out["oracle_polish"] = tot
    if cfg != 1 and not stall:
        est, it, status = newton_correction_estimate(s, d, x, p)
        out["err_estimate_newton_correction"] = est
        out["estimate_mg_pcg"] = {"iters": it, "status": status}
    out.pop("steps")
    print("FULLCONFIG " + json.dumps(out), flush=True)
    if not stall:
        assert out["converged"] and out["oracle_rel_residual"] < 1e-8
        assert out["err_estimate_newton_correction"] < 1e-6
        if "err_vs_fp64_oracle" in out:
            assert out["err_vs_fp64_oracle"] < 1e-6
            # the estimate agrees with the measured error within a factor of 3
            # (where that error is above the polish's own resolution)
            if cfg != 1 and out["err_vs_fp64_oracle"] > 1e-11:
                e, m = out["err_estimate_newton_correction"], out["err_vs_fp64_oracle"]
                assert m / 3 <= e <= 3 * m + 1e-12, (e, m)

"""The five beam workloads of this solver's evaluation, built from the mesh
generator, and a runner that solves one on the device and reports iterations,
timings and FP64 residuals.

    python -m solidfem_amd.configs 1 [jacobi|multigrid|multigrid_learned|learned|none] [nopredictor|secant] [polar]

(wall times are measured after the same solve on a 16x4x4 grid, which loads
every kernel's code object first)

All beams span [0, 1] x [0, 0.25] x [0, 0.25] m with the x = 0 face clamped
and 6 tets per hex (the Models/box2.tet pattern, see mesh.py).

1  10x3x3 (540 tets, 176 nodes), LINEAR E 1e6 Pa, nu 0.45, rho 1070,
   gravity -9.8 z, static: the reference's own Newton-CG semantics
   (10 steps, CG eps 1e-3f) and the converged solve to 1e-8.
2  64x16x16 (98,304 tets, 18,785 nodes), NEO_HOOKEAN E 1e6, nu 0.45, gravity
   plus a follower pressure of 1e4 Pa on the x = 1 face (the reference's
   SetBoundaryConditionsSurface load), 10 quasi-static load steps to 1e-8.
3  128x32x32 (786,432 tets, 140,481 nodes), COROTATIONAL (the reference's QR
   rotation and tangent), per-element E = 10^U(5,7), nu = U(0.30,0.48) from
   seed 0, 20 load steps of a dead tip torsion t = tau e_x x (p - c),
   tau = 1e5 Pa/m about the face centroid c; Jacobi-PCG.  `polar` runs it
   with COROTATIONAL_POLAR (the polar rotation and its consistent tangent):
   the reference's tangent is not the derivative of its force, and Newton
   stalls under the torsion's large rotations.
4  learned preconditioner evaluation beam: 128x32x32 NEO_HOOKEAN as bench.py.
5  256x64x64 (6,291,456 tets, 1,085,825 nodes), nodes jittered U(-0.1,0.1)
   cell (boundary nodes in-plane), NEO_HOOKEAN E 1e6, nu 0.495, gravity plus a
   dead tip traction of 1e4 Pa along -z, 4 load steps.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import capi
from .mesh import BeamMesh, beam_hierarchy, torsion_traction, traction_nodal_forces, uniform01

DESCRIPTIONS = {
    1: "10x3x3 linear cantilever under gravity, static",
    2: "64x16x16 Neo-Hookean, gravity + 1e4 Pa follower tip pressure, 10 load steps",
    3: "128x32x32 corotational, per-element E/nu, 20 steps of tip torsion",
    4: "128x32x32 Neo-Hookean, gravity + 1e4 Pa dead tip traction, 10 load steps",
    5: "256x64x64 jittered Neo-Hookean nu=0.495, gravity + 1e4 Pa dead tip traction, 4 load steps",
}


def build(cfg, grid=None, material=None):
    """-> dict(mesh, fixed, material, young, poisson, young_e, poisson_e,
    gravity, tris, pressure, fdead, load_steps); `material` replaces the
    workload's material model"""
    g = {1: (10, 3, 3), 2: (64, 16, 16), 3: (128, 32, 32), 4: (128, 32, 32), 5: (256, 64, 64)}[cfg]
    grid = grid or g
    m = BeamMesh(*grid, jitter=0.1 if cfg == 5 else 0.0, seed=5 if cfg == 5 else 0)
    d = dict(mesh=m, fixed=m.fixed_mask(), young=1e6, poisson=0.45, young_e=None, poisson_e=None,
             gravity=(0.0, 0.0, -9.8), tris=None, pressure=0.0, fdead=None, load_steps=1)
    tris = m.tip_triangles()
    if cfg == 1:
        d.update(material=capi.LINEAR)
    elif cfg == 2:
        d.update(material=capi.NEO_HOOKEAN, tris=tris, pressure=1e4, load_steps=10)
    elif cfg == 3:
        d.update(material=capi.COROTATIONAL, gravity=(0.0, 0.0, 0.0), load_steps=20,
                 young_e=10 ** (5 + 2 * uniform01(0, 1, m.ne)), poisson_e=0.30 + 0.18 * uniform01(0, 2, m.ne),
                 fdead=traction_nodal_forces(m.X, tris, torsion_traction((1.0, 0.125, 0.125), 1e5)))
    elif cfg == 4:
        d.update(material=capi.NEO_HOOKEAN, load_steps=10,
                 fdead=traction_nodal_forces(m.X, tris, lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1))))
    else:
        d.update(material=capi.NEO_HOOKEAN, poisson=0.495, load_steps=4,
                 fdead=traction_nodal_forces(m.X, tris, lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1))))
    if material is not None:
        d["material"] = material
    return d


def solver(d):
    m = d["mesh"]
    s = capi.Solver(m.X, m.tets, d["fixed"], d["material"], d["young"], d["poisson"], 1070.0,
                    gravity=d["gravity"], young_e=d["young_e"], poisson_e=d["poisson_e"])
    if d["tris"] is not None and d["pressure"]:
        s.set_pressure_surface(d["tris"], d["pressure"])
    if d["fdead"] is not None:
        s.set_external_forces(d["fdead"])
    return s


PRECONDS = {"none": capi.PC_NONE, "jacobi": capi.PC_JACOBI, "learned": capi.PC_LEARNED,
            "multigrid": capi.PC_MULTIGRID, "multigrid_learned": capi.PC_MULTIGRID_LEARNED}


def run(cfg, precond=capi.PC_JACOBI, tol=1e-8, grid=None, max_newton=50, log=None, predictor=2, material=None,
        cg_rtol=0.0):
    """Solves workload `cfg`; `log(str)` (optional) gets one line per load step.
    The multigrid preconditioners get the beam's nested hierarchy, the learned
    ones the shipped weights."""
    d = build(cfg, grid, material)
    s = solver(d)
    if precond in (capi.PC_MULTIGRID, capi.PC_MULTIGRID_LEARNED):
        s.set_multigrid(beam_hierarchy(d["mesh"], young_e=d["young_e"], poisson_e=d["poisson_e"]))
    if precond in (capi.PC_LEARNED, capi.PC_MULTIGRID_LEARNED):
        from . import learned
        smoother = precond == capi.PC_MULTIGRID_LEARNED and os.path.exists(learned.SMOOTHER_WEIGHTS)
        net = learned.load(smoother=smoother)
        net.attach(s)
        s._net = net   # keep the network alive with the solver
    steps = d["load_steps"]
    p = capi.solve_params(tol=tol, precond=precond, cg_rtol=cg_rtol, outer=max_newton, inner=200000, check_every=16,
                          sim_type=capi.ST_QUASI_STATIC if steps > 1 else capi.ST_STATIC, force_step=1.0 / steps,
                          predictor=predictor)
    out = {"config": cfg, "description": DESCRIPTIONS[cfg], "tets": d["mesh"].ne, "nodes": d["mesh"].nn,
           "material": d["material"], "steps": []}
    t0 = time.perf_counter()
    for k in range(steps):
        st = s.step(p, allow_notconv=True)
        out["steps"].append({k: st[k] for k in ("newton_iters", "cg_iters", "fallbacks", "rejects", "converged", "res",
                                                "ext_norm",
                                                "seconds")})
        if log:
            log(f"workload {cfg} load step {k + 1}/{steps}: newton {st['newton_iters']} cg {st['cg_iters']} "
                f"fallbacks {st['fallbacks']} rejected predictions {st['rejects']} rel residual {st['res'] / st['ext_norm']:.2e} ({st['seconds']:.2f} s)")
    out["wall_s"] = time.perf_counter() - t0
    out["newton_iters"] = sum(x["newton_iters"] for x in out["steps"])
    out["cg_iters"] = sum(x["cg_iters"] for x in out["steps"])
    out["fallbacks"] = sum(x["fallbacks"] for x in out["steps"])
    out["rejects"] = sum(x["rejects"] for x in out["steps"])
    out["converged"] = all(x["converged"] for x in out["steps"])
    out["final_rel_residual"] = out["steps"][-1]["res"] / out["steps"][-1]["ext_norm"]
    return out, d, s


if __name__ == "__main__":
    cfg = int(sys.argv[1])
    pc = sys.argv[2] if len(sys.argv) > 2 else "jacobi"
    pred = 0 if "nopredictor" in sys.argv[3:] else (1 if "secant" in sys.argv[3:] else 2)
    mat = capi.COROTATIONAL_POLAR if "polar" in sys.argv[3:] else None
    if cfg != 1:
        # the same solve on a 16x4x4 grid first: the first launch of each
        # kernel loads its code object, which would be timed with load step 1
        run(cfg, PRECONDS[pc], grid=(16, 4, 4), predictor=pred, material=mat)
    out, d, s = run(cfg, PRECONDS[pc], log=lambda line: print(line, file=sys.stderr, flush=True), predictor=pred,
                    material=mat)
    out["precond"] = pc
    out["predictor"] = pred
    print(json.dumps(out), flush=True)

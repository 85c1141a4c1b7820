"""Benchmark: Newton-CG solve of the 786,432-tet Neo-Hookean beam on one MI355X.

Workload (one timed step = one complete quasi-static solve from rest):
  128x32x32 hex grid, 6 tets/hex (786,432 tets, 140,481 nodes) over
  [0,1]x[0,0.25]x[0,0.25] m, x = 0 face clamped; Neo-Hookean E = 1e6 Pa,
  nu = 0.45, rho = 1070 kg/m^3; gravity -9.8 m/s^2 along z plus a dead tip
  traction of 1e4 Pa along -z on the x = 1 face; 10 load steps, each Newton-PCG
  to ||r||_2 <= 1e-8 ||r_ext||_2 in FP64 (FP32 element kernels and CG operator).

Prints ONE JSON line.  `value` = element evaluations per second over the timed
solves (Hessian-vector + internal-force element evaluations / wall time), with
the mesh and all state resident in HBM before the timed region starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from solidfem_amd import capi  # noqa: E402
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, traction_nodal_forces  # noqa: E402

HBM_PEAK_GBS = 8000.0
# every TIME_EVERY-th Hessian-vector launch of the timed solves is bracketed by
# HIP events on the solver's stream (the roofline's average launch duration)
TIME_EVERY = 16
COUNTERS = os.path.join(ROOT, "profiles", "hessvec_nh786k_counters.md")


def measured_traffic(grid, key):
    """HBM bytes per launch of a Hessian-vector kernel on the 128x32x32 beam
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (its line
    `key: N`); None for other grids or without it."""
    if grid != (128, 32, 32) or not os.path.exists(COUNTERS):
        return None
    with open(COUNTERS) as f:
        for line in f:
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    return None


def workload(grid=(128, 32, 32)):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    tris = m.tip_triangles()
    fdead = traction_nodal_forces(m.X, tris, lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1)))
    return m, fixed, fdead


def hessvec_bytes(ne, nn, per_element=False, direction_update=False):
    """Algorithmic bytes of one Hessian-vector apply: per element 16 B node ids +
    36 B Dm^-1 + 4 B volume (+ 8 B mu, lambda when per element), per node 36 B
    (read u, read d, write Kd), FP32.  direction_update: the CG loop's element
    kernel, which also forms d = z + beta p_old (per node 12 B more read, z,
    and 12 B written, the new p)."""
    return ne * (16 + 36 + 4 + (8 if per_element else 0)) + nn * (36 + (24 if direction_update else 0))


def host_info():
    """nproc, the threads the CPU legs use (OMP_NUM_THREADS where the host sets
    it, e.g. the CPU share of a GPU box, else every core) and lscpu's model"""
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc))
    info = {"nproc": nproc, "threads": threads}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except Exception:   # lscpu absent: nproc alone
        pass
    return info


def _rate(fn, budget_s):
    """calls of fn per second over about budget_s seconds (after one warm call)"""
    fn()
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        dt = time.perf_counter() - t0
        if dt > budget_s or n >= 1000:
            return n / dt, n, dt


def cpu_baseline(m, fixed, fdead, x, gpu_runs, budget_s=8.0):
    """The CPU path on this host's cores, on bounded samples of the same
    workload: Hessian-vector products at the solved state through the
    compiled reference's own MatrixVectorMultiply (oracle/_ref, built from the
    reference sources; OpenMP element loop, serial scatter) when that library
    is present, and through the FP64 restatement with the same loop structure
    (oracle/sfem_oracle.c) always; plus the restatement's Jacobi-PCG iteration
    and residual, which extrapolate its converged Newton-CG over the GPU
    solve's iteration counts (labelled as extrapolated)."""
    from oracle import oracle as O
    host = host_info()
    os.environ.setdefault("OMP_NUM_THREADS", str(host["threads"]))
    threads = host["threads"]
    d = np.random.default_rng(0).standard_normal(m.X.shape)
    p = O.OracleProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8),
                        fext=fdead, nthreads=threads, incidence=True)
    port_rate, n_port, dt_port = _rate(lambda: p.matvec(x, d), budget_s)
    out = {"unit": "elem-evals/s", "cores": threads, "host": host,
           "port_value": port_rate * m.ne,
           "port_sample": f"{n_port} Hessian-vector products of the {m.ne:,}-tet NH beam at its solved state "
                          f"(oracle/sfem_oracle.c orc_matvec, {threads} OpenMP threads, {dt_port:.1f} s)"}
    if O.ref_available():
        ref = O.RefProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, -9.8, zup=True)
        ref.set_positions(x)
        ref_rate, n_ref, dt_ref = _rate(lambda: ref.matvec(d), budget_s)
        out.update(value=ref_rate * m.ne, kind="reference",
                   sample=f"{n_ref} Hessian-vector products of the {m.ne:,}-tet NH beam at its solved state "
                          f"(FemPhysicsMatrixFree::MatrixVectorMultiply compiled from the reference sources, "
                          f"{threads} OpenMP threads, {dt_ref:.1f} s)")
    else:
        out.update(value=out["port_value"], kind="port", sample=out["port_sample"])
    # the restatement's converged Jacobi-PCG: one PCG iteration (operator,
    # block-Jacobi apply, vector updates) and one FP64 residual, timed
    B = p.block_diag(x)
    Binv = np.linalg.inv(B).reshape(-1, 9)
    b = p.gradients(x, 1.0)
    cg_rate, n_cg, dt_cg = _rate(lambda: p.cg(x, b, 10, 1e-30, Binv), budget_s / 2)
    res_rate, n_res, _ = _rate(lambda: p.gradients(x, 1.0), budget_s / 4)
    cg_total = sum(r["cg_iters"] for r in gpu_runs) / len(gpu_runs)
    res_total = sum(r["force_evals"] for r in gpu_runs) / len(gpu_runs) / m.ne
    t_iter, t_res = 1.0 / (10 * cg_rate), 1.0 / res_rate
    out["newton_cg_extrapolated_s"] = cg_total * t_iter + res_total * t_res
    out["newton_cg_extrapolation"] = (
        f"EXTRAPOLATED: the GPU Jacobi-PCG solve's {cg_total:.0f} CG iterations x {1e3 * t_iter:.2f} ms "
        f"(oracle block-Jacobi PCG iteration, {threads} threads) + {res_total:.0f} FP64 residuals x "
        f"{1e3 * t_res:.2f} ms; the reference's own SolveNewtonCG (unpreconditioned CG, fixed 1e-3 "
        f"forcing, no residual test) does not reach 1e-8 on this beam")
    return out


def cpu_full_newton(grid, load_steps, threads):
    """The FP64 restatement's converged Newton with block-Jacobi PCG, run in
    full on the CPU (minutes): the CPU counterpart of the GPU Jacobi-PCG leg"""
    from oracle import oracle as O
    m, fixed, fdead = workload(grid)
    p = O.OracleProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0, 0, -9.8),
                        fext=fdead, nthreads=threads, incidence=True)
    x = m.X.copy()
    tot = {"newton": 0, "cg": 0}
    t0 = time.perf_counter()
    for k in range(load_steps):
        x, st = p.newton_solve(x, (k + 1) / load_steps, tol=1e-8, cg_rtol=0.1, max_newton=60, max_cg=1000000)
        tot["newton"] += st["newton"]
        tot["cg"] += st["cg"]
        print(f"cpu load step {k + 1}/{load_steps}: newton {st['newton']} cg {st['cg']} "
              f"rel residual {st['res'] / st['ext']:.2e} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    return {"grid": "x".join(map(str, grid)), "tets": m.ne, "load_steps": load_steps, "threads": threads,
            "newton_cg_wall_s": time.perf_counter() - t0, **tot, "host": host_info()}


MATERIALS = {"linear": capi.LINEAR, "stvk": capi.STVK, "corotational": capi.COROTATIONAL,
             "neo-hookean": capi.NEO_HOOKEAN, "corotational-polar": capi.COROTATIONAL_POLAR}


def hessvec_only(s, m, args):
    """Kernel-level timing at a smooth deformed state (bending-like field), no
    solve: the run that hardware-counter passes profile."""
    x = m.X.copy()
    xi = m.X[:, 0]
    x[:, 2] += -0.3 * xi ** 2
    x[:, 0] += 0.01 * np.sin(3 * m.X[:, 1]) * xi
    s.set_positions(x)
    ms_elem, ms_gather = s.profile_hessvec(args.hessvec_only)
    info = s.info()
    bytes_alg = hessvec_bytes(info["ne"], info["nn"], per_element=args.per_element_moduli)
    out = {"kernel": f"k_elem_block<0,{args.material.upper().replace('-', '_')},256>", "material": args.material,
           "per_element_moduli": args.per_element_moduli, "reps": args.hessvec_only, "ms_elem": ms_elem,
           "ms_gather": ms_gather, "bytes_alg": bytes_alg, "gbs_alg": bytes_alg / (ms_elem * 1e-3) / 1e9}
    if args.calib_mib:
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        n = args.calib_mib << 20
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
        assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0        # device to device: one copy kernel
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(a)
        hip.hipFree(b)
        out["calib_copy_bytes"] = n
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precond", choices=["all", "both", "jacobi", "none", "learned", "multigrid", "multigrid_learned"],
                    default="all",
                    help="all: the Jacobi-PCG solve is `value`; the learned-PCG solves (learned blocks alone, and "
                         "as the finest smoother of the multigrid preconditioner) and the multigrid-PCG solve with "
                         "block-Jacobi smoothers are reported beside it; both: Jacobi and learned blocks only")
    ap.add_argument("--load-steps", type=int, default=10)
    ap.add_argument("--grid", default="128x32x32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--hessvec-only", type=int, default=0, metavar="REPS",
                    help="only time REPS Hessian-vector launches at a deformed state (for counter runs)")
    ap.add_argument("--material", choices=list(MATERIALS), default="neo-hookean",
                    help="with --hessvec-only: the material whose Hessian-vector kernel is timed")
    ap.add_argument("--calib-mib", type=int, default=0,
                    help="with --hessvec-only: also run one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--per-element-moduli", action="store_true",
                    help="with --hessvec-only: E and nu per element, drawn as in the per-element-material workload "
                         "(E log-uniform in [1e5, 1e7] Pa, nu uniform in [0.30, 0.48])")
    ap.add_argument("--tangent-cache", type=int, choices=[0, 1], default=1,
                    help="Neo-Hookean tangent state cached per Newton step (1, the default) or rebuilt in every "
                         "Hessian-vector product (0)")
    ap.add_argument("--cpu-full-newton", action="store_true",
                    help="only run the CPU FP64 restatement's converged Jacobi-PCG Newton on --grid in full "
                         "(no GPU), and print its wall time as one JSON line")
    args = ap.parse_args()
    if args.cpu_full_newton:
        print(json.dumps(cpu_full_newton(tuple(int(v) for v in args.grid.split("x")), args.load_steps,
                                         host_info()["threads"])), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    grid = tuple(int(v) for v in args.grid.split("x"))
    m, fixed, fdead = workload(grid)
    if args.material != "neo-hookean" and not args.hessvec_only:
        ap.error("--material applies to --hessvec-only; the solve legs run the Neo-Hookean workload")
    per_e = {}
    if args.per_element_moduli:
        if not args.hessvec_only:
            ap.error("--per-element-moduli applies to --hessvec-only")
        from solidfem_amd.mesh import uniform01
        per_e = dict(young_e=10 ** (5 + 2 * uniform01(0, 1, m.ne)), poisson_e=0.30 + 0.18 * uniform01(0, 2, m.ne))
    s = capi.Solver(m.X, m.tets, fixed, MATERIALS[args.material], 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8), **per_e)
    s.set_external_forces(fdead)
    s.set_tangent_cache(args.tangent_cache)
    if args.hessvec_only:
        hessvec_only(s, m, args)
        return
    def leg(pc):
        """warmup + timed solves with preconditioner pc -> (wall s, runs, final positions)"""
        params = capi.solve_params(tol=1e-8, precond=pc, cg_rtol=0.0, sim_type=capi.ST_QUASI_STATIC,
                                   force_step=1.0 / args.load_steps, inner=100000, check_every=16)

        def one_solve():
            s.reset()
            tot = {"newton_iters": 0, "cg_iters": 0, "hessvec_evals": 0, "force_evals": 0, "converged": 1,
                   "fallbacks": 0, "rejects": 0}
            last = None
            for _ in range(args.load_steps):
                st = s.step(params)
                for k in ("newton_iters", "cg_iters", "hessvec_evals", "force_evals", "fallbacks", "rejects"):
                    tot[k] += st[k]
                tot["converged"] &= st["converged"]
                last = st
            tot["res"], tot["ext_norm"] = last["res"], last["ext_norm"]
            return tot

        for _ in range(args.warmup):
            one_solve()
        s.time_hessvec(TIME_EVERY)
        t0 = time.perf_counter()
        runs = [one_solve() for _ in range(args.steps)]
        dt = time.perf_counter() - t0
        ms, n = s.time_hessvec(0)
        return dt, runs, s.positions(), (ms, n)

    main_pc = {"jacobi": capi.PC_JACOBI, "both": capi.PC_JACOBI, "all": capi.PC_JACOBI, "none": capi.PC_NONE,
               "learned": capi.PC_LEARNED, "multigrid": capi.PC_MULTIGRID,
               "multigrid_learned": capi.PC_MULTIGRID_LEARNED}[args.precond]
    net = None
    if args.precond in ("learned", "multigrid_learned", "both", "all"):
        from solidfem_amd import learned
        if args.precond == "multigrid_learned" and os.path.exists(learned.SMOOTHER_WEIGHTS):
            net = learned.load(smoother=True)
            net.attach(s)
        elif os.path.exists(learned.WEIGHTS):
            net = learned.load()
            net.attach(s)
        elif args.precond in ("learned", "multigrid_learned"):
            raise SystemExit(f"no trained weights at {learned.WEIGHTS}")
    hierarchy = None
    if args.precond in ("multigrid", "multigrid_learned", "all"):
        levels = beam_hierarchy(m)
        s.set_multigrid(levels)
        # grid sizes finest to coarsest (each level halves the hex grid); the last is solved directly
        grids = [tuple(g // 2 ** k for g in grid) for k in range(len(levels) + 1)]
        hierarchy = (" > ".join("x".join(map(str, g)) for g in grids)
                     + f" (coarsest solved directly, {3 * levels[-1][0].shape[0]} dofs)")
    dt, runs, x, (ms_elem, n_timed) = leg(main_pc)
    evals = sum(r["hessvec_evals"] + r["force_evals"] for r in runs)
    cg = sum(r["cg_iters"] for r in runs)

    def side_leg(pc):
        dtl, runs_l, xl, _ = leg(pc)
        ev_l = sum(r["hessvec_evals"] + r["force_evals"] for r in runs_l)
        return {
            "newton_solve_s": dtl / args.steps, "value": ev_l / dtl, "unit": "elem-evals/s",
            "newton_iters": runs_l[-1]["newton_iters"], "cg_iters": runs_l[-1]["cg_iters"],
            "fallbacks": runs_l[-1]["fallbacks"], "converged": bool(all(r["converged"] for r in runs_l)),
            "final_rel_residual": runs_l[-1]["res"] / runs_l[-1]["ext_norm"],
            "speedup_vs_jacobi": dt / dtl,
            "displacement_rel_diff_vs_jacobi": float(np.linalg.norm(xl - x) / np.linalg.norm(x - m.X)),
            **({"hierarchy": hierarchy} if pc in (capi.PC_MULTIGRID, capi.PC_MULTIGRID_LEARNED) else {}),
        }

    extra = {}
    if args.precond in ("both", "all"):
        if net is None:
            extra["learned_pcg"] = {"skipped": "no trained weights in the package"}
        else:
            extra["learned_pcg"] = {"block": side_leg(capi.PC_LEARNED)}
            if args.precond == "all":
                # the network trained as the smoother (white-noise pairs), when shipped
                if os.path.exists(learned.SMOOTHER_WEIGHTS):
                    smoother = learned.load(smoother=True)
                    smoother.attach(s)
                extra["learned_pcg"]["multigrid"] = side_leg(capi.PC_MULTIGRID_LEARNED)
        if args.precond == "all":
            extra["multigrid_pcg"] = side_leg(capi.PC_MULTIGRID)

    info = s.info()
    # the timed launches of the Jacobi-PCG loop are the element kernel with the
    # direction update (every iteration but the first of each solve)
    bytes_alg = hessvec_bytes(info["ne"], info["nn"], direction_update=n_timed > 0)
    timing = "in-solve"
    if n_timed == 0:
        # the multigrid loops launch the kernel without events: time it back to back
        ms_elem, _ = s.profile_hessvec(200)
        timing = "back-to-back"
        bytes_alg = hessvec_bytes(info["ne"], info["nn"])
    achieved = bytes_alg / (ms_elem * 1e-3) / 1e9
    # the whole stand-alone apply (element blocks + shared-node completion),
    # back to back at the solution state
    ms_e2, ms_g2 = s.profile_hessvec(200)
    bytes_apply = hessvec_bytes(info["ne"], info["nn"])
    line = {
        "metric": "Newton-CG time to 1e-8 rel. residual (learned-PCG vs Jacobi-PCG); elem-evals/s",
        "value": evals / dt,
        "unit": "elem-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 element kernels + f64 residual/dots",
        "data": "synthetic",
        "config": {"workload": f"nh_beam_{info['ne']}tet_{args.load_steps}loadsteps",
                   "grid": args.grid, "tets": info["ne"], "nodes": info["nn"], "material": "neo-hookean",
                   "E": 1e6, "nu": 0.45, "loads": "gravity -z + dead tip traction 1e4 Pa -z",
                   "load_steps": args.load_steps, "tol": 1e-8,
                   "precond": "jacobi" if args.precond in ("both", "all") else args.precond,
                   "tangent_cache": args.tangent_cache, "parallelism": "none"},
        "newton_solve_s": dt / args.steps,
        "newton_iters": runs[-1]["newton_iters"],
        "cg_iters": runs[-1]["cg_iters"],
        "cg_iters_per_s": cg / dt,
        "converged": bool(all(r["converged"] for r in runs)),
        "final_rel_residual": runs[-1]["res"] / runs[-1]["ext_norm"],
        "tip_deflection_m": float((x - m.X)[:, 2].min()),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(grid, "traffic_bytes_per_launch_cg_lin" if args.tangent_cache
                                                 else "traffic_bytes_per_launch_cg"),
                     "traffic_source": "profile-derived: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel in a "
                                       "one-load-step Jacobi-PCG solve (profiles/hessvec_nh786k_counters.md), not "
                                       "measured in this run",
                     "kernel": (("k_elem_block<0,NEO_HOOKEAN_LIN,256,PF> (Hessian-vector element blocks on the "
                                 "Neo-Hookean tangent state cached per Newton step, with the CG direction update)"
                                 if args.tangent_cache else
                                 "k_elem_block<0,NEO_HOOKEAN,256,PF> (Hessian-vector element blocks with the CG "
                                 "direction update)") if n_timed else
                                "k_elem_block<0,NEO_HOOKEAN,256> (Hessian-vector element blocks)"),
                     "ms_per_launch": ms_elem, "timed_launches": n_timed, "timing": timing,
                     "bytes_per_launch": bytes_alg,
                     "whole_apply": {"ms_elem": ms_e2, "ms_shared_gather": ms_g2, "bytes": bytes_apply,
                                     "achieved_gbs": bytes_apply / ((ms_e2 + ms_g2) * 1e-3) / 1e9,
                                     "frac": bytes_apply / ((ms_e2 + ms_g2) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                     "timing": "back-to-back at the final state"}},
    }
    line.update(extra)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(m, fixed, fdead, x, runs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

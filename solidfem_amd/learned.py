"""Training of the learned block preconditioner on pairs the solver generates,
and the trained weights it ships with.

    python -m solidfem_amd.learned train [--beams 24] [--epochs 30] [--out FILE]
    python -m solidfem_amd.learned train-smoother [--rough 84]

The weights in weights/learned_precond_nh.npz were written by `train`
(2,016 pairs on 24 beams, 30 epochs, 47 s on one MI355X); bench.py loads them
for its learned-PCG leg (the learned blocks as the preconditioner).  Those in
weights/learned_smoother_nh.npz were written by `train-smoother` (84
white-noise pairs per beam, the same 24 beams and schedule, 42 s): the
network as the finest-level smoother of the multigrid preconditioner, whose
job is the high-frequency error (held-out one-step loss 0.108 vs 0.123 for
block Jacobi).

Training data (the network, loss and device kernels are those of
sfem_gnn_* in include/sfem_capi.h): seeded random Neo-Hookean beams from
32x8x8 to 128x32x32 hex cells (6 tets each), half with uniform materials
(E = 1e6 * 10^U(-0.5, 0.5) Pa, nu = U(0.30, 0.495)) and half with per-element
fields (E log-uniform in [1e5, 1e7] Pa, nu = U(0.30, 0.48)), every other one
with nodes jittered by +-0.1 cell.  Each beam is first driven by Newton to a
random fraction of gravity plus a random tip traction (|t| <= 1e4 Pa), so the
tangent is that of a deformed state, as inside a Newton solve; at that state
the solver collects (r_k, r_k . K^-1 r_k) pairs along block-Jacobi PCG runs
on random loads.  Training is Adam on the one-step energy-error loss,
alternating over the beams.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi
from .mesh import BeamMesh, traction_nodal_forces, uniform01

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "learned_precond_nh.npz")
# the network trained as the finest-level smoother of the multigrid preconditioner
SMOOTHER_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "learned_smoother_nh.npz")
GRIDS = [(32, 8, 8), (48, 12, 12), (64, 16, 16), (96, 24, 24), (128, 32, 32)]


def random_beam(seed, grid):
    """-> (mesh, fixed, solver) of training beam `seed`, at its deformed state."""
    rng = np.random.default_rng(1000 + seed)
    m = BeamMesh(*grid, jitter=0.1 if seed % 2 else 0.0, seed=seed)
    fixed = m.fixed_mask()
    if seed % 4 < 2:
        E = 1e6 * 10 ** rng.uniform(-0.5, 0.5)
        nu = rng.uniform(0.30, 0.495)
        kw = {}
    else:
        E, nu = 1e6, 0.45
        kw = dict(young_e=10 ** (5 + 2 * uniform01(seed, 1, m.ne)), poisson_e=0.30 + 0.18 * uniform01(seed, 2, m.ne))
    s = capi.Solver(m.X, m.tets, fixed, capi.NEO_HOOKEAN, E, nu, 1070.0, gravity=(0.0, 0.0, -9.8), **kw)
    t = rng.standard_normal(3)
    t *= rng.uniform(0.0, 1e4) / np.linalg.norm(t)
    s.set_external_forces(traction_nodal_forces(m.X, m.tip_triangles(), lambda p: np.tile(t, (len(p), 1))))
    frac = rng.uniform(0.0, 1.0)
    steps = 3
    p = capi.solve_params(tol=1e-6, precond=capi.PC_JACOBI, cg_rtol=0.1, outer=30, inner=100000, check_every=32,
                          sim_type=capi.ST_QUASI_STATIC, force_step=frac / steps)
    for _ in range(steps):
        s.step(p, allow_notconv=True)
    return m, fixed, s


def random_load(m, fixed, rng):
    """a random tip traction plus a random nodal field, zero on fixed nodes"""
    t = rng.standard_normal(3) * 1e4
    f = traction_nodal_forces(m.X, m.tip_triangles(), lambda p: np.tile(t, (len(p), 1)))
    f += rng.standard_normal(m.X.shape) * rng.uniform(0.0, 0.5) * np.abs(f).max()
    return f * (1 - fixed[:, None])


def collect(seed, grid, loads=3, nsnap=28, rough=0):
    """pairs on training beam `seed`: PCG snapshots under `loads` random
    loads, or (rough > 0) that many white-noise errors (smoother training)"""
    m, fixed, s = random_beam(seed, grid)
    if rough:
        capi.LearnedPreconditioner.collect_rough(s, rough, seed=3000 + seed)
        return s
    rng = np.random.default_rng(2000 + seed)
    # snapshot spacing grows with the beam: the run spans most of a solve
    every = max(4, grid[0] // 2)
    for _ in range(loads):
        capi.LearnedPreconditioner.collect(s, random_load(m, fixed, rng), nsnap=nsnap, snap_every=every, rtol=1e-6)
    return s


def train(beams=24, epochs=30, steps=20, batch=8, lr=2e-3, seed=1, grids=GRIDS, held_out=3, log=print, rough=0):
    """-> (trained LearnedPreconditioner, summary dict); rough > 0 trains the
    smoother on that many white-noise pairs per beam"""
    t0 = time.perf_counter()
    train_set = [collect(k, grids[k % len(grids)], rough=rough) for k in range(beams)]
    held = [collect(10000 + k, grids[(k + 2) % len(grids)], rough=rough) for k in range(held_out)]
    npairs = sum(capi.LearnedPreconditioner.num_pairs(s) for s in train_set)
    log(f"collected {npairs} pairs on {beams} beams in {time.perf_counter() - t0:.1f} s")
    net = capi.LearnedPreconditioner(seed=seed)
    hist = []
    for ep in range(epochs):
        for k, s in enumerate(train_set):
            net.train(s, steps, batch=batch, lr=lr * (0.5 if ep >= epochs * 2 // 3 else 1.0), seed=ep * 1000 + k)
        ev = [net.evaluate(s) for s in held]
        ll, lj = float(np.mean([a for a, _ in ev])), float(np.mean([b for _, b in ev]))
        hist.append((ll, lj))
        log(f"epoch {ep}: held-out loss learned {ll:.5f} jacobi {lj:.5f}")
    return net, {"pairs": npairs, "beams": beams, "epochs": epochs, "history": hist,
                 "seconds": time.perf_counter() - t0}


def save(net, path=WEIGHTS, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.params(), meta=json.dumps(meta or {}))


def load(path=None, seed=1, smoother=False):
    """-> LearnedPreconditioner with the weights of `path` (default: the
    shipped preconditioner weights, or the smoother weights)"""
    path = path or (SMOOTHER_WEIGHTS if smoother else WEIGHTS)
    return _load(path, seed)


def _load(path, seed=1):
    """-> LearnedPreconditioner with the weights of `path`"""
    if not os.path.exists(path):
        raise FileNotFoundError(f"no trained weights at {path}; run python -m solidfem_amd.learned train")
    with np.load(path, allow_pickle=False) as z:
        w = z["params"]
    net = capi.LearnedPreconditioner(seed=seed)
    net.set_params(w)
    return net


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "train-smoother"])
    ap.add_argument("--beams", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--rough", type=int, default=84, help="white-noise pairs per beam (train-smoother)")
    a = ap.parse_args()
    smoother = a.cmd == "train-smoother"
    net, meta = train(a.beams, a.epochs, a.steps, lr=a.lr, log=lambda s: print(s, flush=True),
                      rough=a.rough if smoother else 0)
    save(net, a.out or (SMOOTHER_WEIGHTS if smoother else WEIGHTS), meta)
    print(json.dumps({k: v for k, v in meta.items() if k != "history"} | {"final": meta["history"][-1]}), flush=True)


if __name__ == "__main__":
    main()

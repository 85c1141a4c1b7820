"""The C++ class's Krylov hooks (MatrixVectorMultiply, ComputeGradients,
DotProduct, UpdatePosAndComputeGradients; FemPhysicsMatrixFree.h:85-89) on
the GPU, driven by the reference's own solver templates instantiated on the
class (oracle/_ref/ref_templates, built from oracle/ref_templates.cpp where
the reference checkout exists): the hooks against the FP64 oracle, the
reference's SolveConjugateGradientMF against the oracle's restatement of the
same CG (Math/IterativeSolvers.h:115-195), and the reference's
NonlinearConjugateGradientMinimizer against a restatement of its loop
(Include/NonlinearCG.h:51-114) on the oracle's FP64 hooks."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd.mesh import BeamMesh

pytestmark = pytest.mark.gpu
DRIVER = O.TEMPLATES_PATH


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def ncg_restated(p, x, frac, free, outer, inner=10, abs_tol=1e-4):
    """NonlinearConjugateGradientMinimizer::Solve restated (float eps and tol
    of 1e-2, Newton-Raphson line search along d)"""
    pos = x.copy()
    eps2 = float(np.float32(1e-2) * np.float32(1e-2))

    def full(v):
        out = np.zeros_like(x)
        out[free] = v
        return out

    r = p.gradients(pos, frac)[free]
    d = r.copy()
    delta = np.sum(r * r)
    delta0 = delta
    for _ in range(outer):
        deltaD = np.sum(d * d)
        for _ in range(inner):
            df = p.matvec(pos, full(d))[free]
            alpha = np.sum(r * d) / np.sum(d * df)
            pos[free] += alpha * d
            r = p.gradients(pos, frac)[free]
            if alpha * alpha * deltaD < eps2:
                break
        delta1 = np.sum(r * r)
        beta = delta1 / delta
        delta = delta1
        if delta < abs_tol * abs_tol or delta <= eps2 * delta0:
            break
        d = r + beta * d
    return pos


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference-template driver not built (no reference checkout)")
def test_reference_templates_on_the_device_class(tmp_path):
    print("case reference CG / NCG templates on FemPhysicsMatrixFree, 10x3x3 NH", flush=True)
    m = BeamMesh(10, 3, 3)
    fixed = m.fixed_mask()
    free = np.flatnonzero(fixed == 0)
    rng = np.random.default_rng(12)
    x = m.X + 0.02 * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    frac = 0.7
    d = rng.standard_normal((free.size, 3))
    b = rng.standard_normal((free.size, 3))
    # 40 CG iterations: the early iterates, which an FP32 operator follows to
    # FP32 accuracy (converging this unpreconditioned CG to its 1e-3 forcing
    # would need ~330 iterations, past where FP32 rounding of the operator
    # lets the recurrence drift from the true residual)
    cg_iters, ncg_outer = 40, 30
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([m.nn, m.ne], dtype=np.int64).tofile(f)
        m.X.astype(np.float64).tofile(f)
        m.tets.astype(np.int32).tofile(f)
        fixed.astype(np.uint8).tofile(f)
        x.astype(np.float64).tofile(f)
        np.array([frac], dtype=np.float64).tofile(f)
        d.tofile(f)
        b.tofile(f)
        np.array([cg_iters, ncg_outer], dtype=np.int32).tofile(f)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    nf3 = 3 * free.size
    off = 0

    def take(n, dt=np.float64):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += n * np.dtype(dt).itemsize
        return a

    Kd, rg, dot = take(nf3), take(nf3), take(1)[0]
    its = int(take(1, np.int32)[0])
    xcg, pos, rn = take(nf3), take(nf3), take(nf3)
    p = O.OracleProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    # the hooks
    dfull = np.zeros_like(m.X)
    dfull[free] = d
    e_kd = rel2(Kd, p.matvec(x, dfull)[free].reshape(-1))
    e_r = np.abs(rg - p.gradients(x, frac)[free].reshape(-1)).max() / np.abs(rg).max()
    print(f"  MatrixVectorMultiply rel err {e_kd:.2e}, ComputeGradients {e_r:.2e}", flush=True)
    assert e_kd < 2e-6 and e_r < 1e-11
    assert abs(dot - np.dot(rg, d.reshape(-1))) <= 1e-13 * abs(dot)
    # the reference's CG on the class vs the oracle's CG (same recurrence, FP64 operator)
    bfull = np.zeros_like(m.X)
    bfull[free] = b
    xo, ko = p.cg(x, bfull, cg_iters, 1e-3)
    e_cg = rel2(xcg, xo[free].reshape(-1))
    res = np.linalg.norm(b.reshape(-1) - p.matvec(x, _full(xcg, free, x))[free].reshape(-1)) / np.linalg.norm(b)
    print(f"  SolveConjugateGradientMF: {its} iterations (oracle {ko}), rel diff {e_cg:.2e}, rel residual {res:.2e}",
          flush=True)
    assert its == ko == cg_iters and e_cg < 1e-4
    # the reference's nonlinear CG on the class vs its loop on the oracle's hooks
    po = ncg_restated(p, x, frac, free, ncg_outer)
    e_ncg = rel2(pos - m.X[free].reshape(-1), (po - m.X)[free].reshape(-1))
    g0, g1 = np.linalg.norm(p.gradients(x, frac)[free]), np.linalg.norm(rn)
    print(f"  NonlinearConjugateGradientMinimizer: rel diff {e_ncg:.2e}, |r| {g0:.3e} -> {g1:.3e}", flush=True)
    assert e_ncg < 1e-3 and g1 < 0.1 * g0


def _full(v, free, x):
    out = np.zeros_like(x)
    out[free] = v.reshape(-1, 3)
    return out

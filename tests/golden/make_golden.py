"""Writes the golden vectors under tests/golden/ from the COMPILED REFERENCE
(oracle/_ref/libsfem_ref.so, built by `make -C oracle ref` from the reference
sources).  Run in a container that has the reference checkout:

    python tests/golden/make_golden.py

Fixtures (npz, float64, original node order):
  ref_cfg1_mat{m}.npz  10x3x3 beam (540 tets, 176 nodes), E=1e6 nu=0.45 rho=1070,
                       gravity -9.8 z, follower pressure 1e4 Pa on the x=L face.
                       x (perturbed positions), d, grad = ComputeGradients(x) at
                       fraction 1, Kd = MatrixVectorMultiply(d) at x,
                       newton_x = positions after SolveNewtonCG (10 outer steps,
                       1000 inner CG iterations at most, eps 1e-3f) from rest.
  ref_box2_nh.npz      Models/box2.tet as the reference loads it (y/z swapped,
                       negative volumes), cantilever.xml material (NH, E 66000,
                       nu 0.45, rho 1070, g -9.8 along y), nodes 0-8 fixed:
                       grad and Kd at a perturbed x.
  ref_tet.npz          the single tetrahedron of PyTester test_tetrahedron
                       (PyTester/PyTester.py:34-60), nodes 0-2 fixed, NH.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from solidfem_amd.mesh import BeamMesh, load_tet1  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    if not O.ref_available():
        O.build_ref()
    rng = np.random.default_rng(20240531)
    m = BeamMesh(10, 3, 3)
    fixed = m.fixed_mask()
    tris = m.tip_triangles()
    for mat in range(4):
        ref = O.RefProblem(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, -9.8, zup=True,
                           tris=tris, pressure=1e4, inner=1000)
        x = m.X + 0.01 * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
        d = rng.standard_normal(m.X.shape)
        ref.set_positions(x)
        ref.set_fraction(1.0)
        grad = ref.gradients()
        kd = ref.matvec(d)
        ref.set_positions(m.X)
        ref.newton_cg(1.0)
        xn = ref.positions()
        np.savez(os.path.join(OUT, f"ref_cfg1_mat{mat}.npz"), x=x, d=d, grad=grad, Kd=kd, newton_x=xn,
                 tets=m.tets, X0=m.X, fixed=fixed, tris=tris)

    X, T = load_tet1(os.path.join(OUT, "box2.tet"))
    fixed = np.zeros(len(X), np.uint8)
    fixed[:9] = 1
    ref = O.RefProblem(X, T, fixed, O.NEO_HOOKEAN, 66000.0, 0.45, 1070.0, -9.8, zup=False)
    # box2 as loaded has negative volumes; the perturbation keeps every F's
    # determinant positive so the Neo-Hookean log stays real
    x = X + 0.002 * rng.standard_normal(X.shape) * (1 - fixed[:, None])
    d = rng.standard_normal(X.shape)
    ref.set_positions(x)
    ref.set_fraction(1.0)
    np.savez(os.path.join(OUT, "ref_box2_nh.npz"), x=x, d=d, grad=ref.gradients(), Kd=ref.matvec(d))

    hw = 0.05
    X = np.array([[0, -hw, hw], [0, -hw, -hw], [0, hw, hw], [2 * hw, -hw, hw]], dtype=np.float64)
    T = np.array([[0, 1, 2, 3]], dtype=np.int32)
    fixed = np.array([1, 1, 1, 0], np.uint8)
    out = {}
    for mat in range(4):
        ref = O.RefProblem(X, T, fixed, mat, 66000.0, 0.45, 1070.0, -9.8, zup=False)
        x = X.copy()
        x[3] += [0.01, -0.02, 0.005]
        d = rng.standard_normal(X.shape)
        ref.set_positions(x)
        ref.set_fraction(1.0)
        out[f"x{mat}"], out[f"d{mat}"] = x, d
        out[f"grad{mat}"], out[f"Kd{mat}"] = ref.gradients(), ref.matvec(d)
    np.savez(os.path.join(OUT, "ref_tet.npz"), **out)
    print("golden vectors written to", OUT)


if __name__ == "__main__":
    main()

"""Edge cases of the device path through the C-ABI: the smallest mesh (one
tetrahedron), a node that no element references, a mesh whose nodes are all
fixed, a solve without loads, and invariance of the converged state under a
renumbering of the nodes (the element blocks, their Morton order and the
shared-node slots all change with it)."""
import numpy as np
import pytest

from oracle import oracle as O
from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, traction_nodal_forces

pytestmark = pytest.mark.gpu


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def one_tet():
    X = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.1]])
    return X, np.array([[0, 1, 2, 3]], dtype=np.int32)


@pytest.mark.parametrize("mat", [capi.LINEAR, capi.STVK, capi.COROTATIONAL, capi.NEO_HOOKEAN, capi.COROTATIONAL_POLAR])
def test_single_tetrahedron(mat):
    print(f"case single tet mat={mat}", flush=True)
    X, T = one_tet()
    fixed = np.array([1, 1, 1, 0], dtype=np.uint8)
    s = capi.Solver(X, T, fixed, mat, 1e6, 0.3, 1000.0, gravity=(0.0, 0.0, -9.8))
    p = O.OracleProblem(X, T, fixed, mat, 1e6, 0.3, 1000.0, gravity=(0.0, 0.0, -9.8))
    x = X.copy()
    x[3] += [0.004, -0.002, 0.003]
    s.set_positions(x)
    d = np.random.default_rng(2).standard_normal(X.shape)
    assert np.abs(s.internal_forces() - p.internal_forces(x)).max() <= 1e-11 * np.abs(p.internal_forces(x)).max()
    assert rel2(s.matvec(d), p.matvec(x, d)) < 2e-6
    if mat != capi.COROTATIONAL:   # the reference corotational tangent does not converge Newton
        s.set_positions(X)
        s.set_load_fraction(1.0)
        st = s.solve(tol=1e-8, precond=capi.PC_JACOBI)
        xo, _ = p.newton_solve(None, 1.0, tol=1e-12, cg_rtol=1e-10)
        assert st["converged"] == 1
        assert rel2(s.positions() - X, xo - X) < 1e-6


def test_node_without_elements_stays_in_place():
    """a node that no element references has no stiffness and no mass: it
    keeps its position, and the rest solves as without it"""
    m = BeamMesh(6, 2, 2)
    fixed = m.fixed_mask()
    Xo = np.vstack([m.X, [[5.0, 5.0, 5.0]]])
    fo = np.append(fixed, 0).astype(np.uint8)
    out = []
    for X, fx in ((m.X, fixed), (Xo, fo)):
        s = capi.Solver(X, m.tets, fx, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
        s.set_load_fraction(1.0)
        st = s.solve(tol=1e-8, precond=capi.PC_JACOBI)
        assert st["converged"] == 1
        out.append(s.positions())
    assert np.array_equal(out[1][-1], Xo[-1])
    assert rel2(out[1][:-1] - m.X, out[0] - m.X) < 1e-6


def test_all_nodes_fixed_and_no_load():
    m = BeamMesh(4, 2, 2)
    s = capi.Solver(m.X, m.tets, np.ones(m.nn, np.uint8), capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0,
                    gravity=(0.0, 0.0, -9.8))
    s.set_load_fraction(1.0)
    st = s.solve(tol=1e-8, precond=capi.PC_JACOBI)
    assert st["converged"] == 1 and np.array_equal(s.positions(), m.X)
    s2 = capi.Solver(m.X, m.tets, m.fixed_mask(), capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0)
    s2.set_load_fraction(1.0)
    st = s2.solve(tol=1e-8, precond=capi.PC_JACOBI)
    assert st["converged"] == 1 and st["cg_iters"] == 0 and np.array_equal(s2.positions(), m.X)


def test_converged_state_is_independent_of_node_numbering():
    m = BeamMesh(16, 4, 4)
    fixed = m.fixed_mask()
    f = traction_nodal_forces(m.X, m.tip_triangles(), lambda p: np.tile([0.0, 0.0, -1e4], (len(p), 1)))
    perm = np.random.default_rng(7).permutation(m.nn)        # new node k is old node perm[k]
    inv = np.argsort(perm)
    res = []
    for X, T, fx, ff in ((m.X, m.tets, fixed, f), (m.X[perm], inv[m.tets].astype(np.int32), fixed[perm], f[perm])):
        s = capi.Solver(X, T, fx, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
        s.set_external_forces(ff)
        s.set_load_fraction(1.0)
        st = s.solve(tol=1e-10, precond=capi.PC_JACOBI)
        assert st["converged"] == 1
        res.append(s.positions())
    assert rel2(res[1][inv] - m.X, res[0] - m.X) < 1e-7

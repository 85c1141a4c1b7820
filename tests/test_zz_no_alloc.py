"""sfem_matvec_noalloc (the allocation-free Hessian-vector product) against
sfem_matvec through the C-ABI: the same kernels on the same inputs, so the
results must be bitwise equal -- on two calls in a row, the second on a
smaller mesh, after which the first mesh is called again -- and a mesh beyond
the static scratch array is refused with an error."""
import numpy as np
import pytest

from solidfem_amd import capi
from solidfem_amd.mesh import BeamMesh, lame, uniform01

pytestmark = pytest.mark.gpu


def solver(grid, mat, per_elem, seed):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    kw = dict(gravity=(0.0, 0.0, -9.8))
    if per_elem:
        E = 10 ** (5 + 2 * uniform01(seed, 1, m.ne))
        nu = 0.30 + 0.18 * uniform01(seed, 2, m.ne)
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, young_e=E, poisson_e=nu, **kw)
    else:
        s = capi.Solver(m.X, m.tets, fixed, mat, 1e6, 0.45, 1070.0, **kw)
    rng = np.random.default_rng(seed)
    x = m.X + 0.05 * min(m.h) * rng.standard_normal(m.X.shape) * (1 - fixed[:, None])
    s.set_positions(x)
    d = rng.standard_normal(m.X.shape)
    return s, d, fixed


def test_noalloc_matches_matvec_on_two_calls_second_smaller():
    print("case noalloc: 24x6x8 NH per-element, then 10x3x3 corotational, then 24x6x8 again", flush=True)
    big, d_big, fixed_big = solver((24, 6, 8), capi.NEO_HOOKEAN, True, 3)
    small, d_small, fixed_small = solver((10, 3, 3), capi.COROTATIONAL, False, 4)
    ref_big = big.matvec(d_big)
    ref_small = small.matvec(d_small)
    q1 = big.matvec_noalloc(d_big)
    q2 = small.matvec_noalloc(d_small)
    assert np.array_equal(q1, ref_big)
    assert np.array_equal(q2, ref_small)
    assert np.all(q1[fixed_big == 1] == 0) and np.all(q2[fixed_small == 1] == 0)
    # the larger mesh after the smaller one: no stale rows of the previous call
    assert np.array_equal(big.matvec_noalloc(d_big), ref_big)


def test_noalloc_node_of_no_element_and_dynamics():
    """a node that no element references (its row is 0 in both entry points)
    and the implicit-dynamics operator M / h^2 + K"""
    print("case noalloc: unreferenced node, implicit dynamics term", flush=True)
    m = BeamMesh(10, 3, 3)
    X = np.vstack([m.X, [[2.0, 2.0, 2.0]]])
    fixed = np.append(m.fixed_mask(), 0)
    s = capi.Solver(X, m.tets, fixed, capi.NEO_HOOKEAN, 1e6, 0.45, 1070.0, gravity=(0.0, 0.0, -9.8))
    d = np.random.default_rng(7).standard_normal(X.shape)
    q = s.matvec_noalloc(d)
    assert np.array_equal(q, s.matvec(d)) and np.all(q[-1] == 0)
    s.step(capi.solve_params(sim_type=capi.ST_IMPLICIT, time_step=0.01, substeps=1, precond=capi.PC_JACOBI))
    q = s.matvec_noalloc(d)
    assert np.array_equal(q, s.matvec(d))


def test_noalloc_refuses_mesh_beyond_scratch():
    """24 B per node of scratch: 1M nodes (one element, the rest referenced by
    none) exceed the 16 MiB array and are refused, with nothing launched"""
    print("case noalloc: 1M-node mesh beyond the scratch limit", flush=True)
    n = 1 << 20
    X = np.zeros((n, 3))
    X[:, 0] = np.arange(n) * 1e-3
    X[1] = (0.0, 1.0, 0.0)
    X[2] = (0.0, 0.0, 1.0)
    X[3] = (0.0, 1.0, 1.0)
    X[0] = (1.0, 0.0, 0.0)
    tets = np.array([[1, 0, 2, 3]], dtype=np.int32)
    vol = np.linalg.det(np.stack([X[0] - X[1], X[2] - X[1], X[3] - X[1]], axis=1))
    if vol < 0:
        tets = np.array([[0, 1, 2, 3]], dtype=np.int32)
    fixed = np.zeros(n, dtype=np.uint8)
    fixed[1] = 1
    s = capi.Solver(X, tets, fixed, capi.LINEAR, 1e6, 0.3, 1000.0)
    with pytest.raises(capi.SfemError) as e:
        s.matvec_noalloc(np.zeros((n, 3)))
    assert e.value.code == -1 and "limit" in str(e.value)
    # the allocating entry point still serves it
    q = s.matvec(np.ones((n, 3)))
    assert q.shape == (n, 3)

"""CPU checks of the nested beam hierarchy the multigrid preconditioner is
built on (solidfem_amd.mesh.beam_hierarchy): level sizes, nesting (coarse
nodes at fine nodes), the prolongation's exactness on fields trilinear in the
grid indices (so constants and linear fields in the unjittered geometry), and
element orientation."""
import numpy as np

import pytest

from solidfem_amd.mesh import BeamMesh, beam_hierarchy, coarse_cell_mean


def dense_p(nf, nc, ptr, col, val):
    P = np.zeros((nf, nc))
    for i in range(nf):
        P[i, col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
    return P


def test_levels_of_the_bench_beam():
    """halving down to the first level the library solves directly (16x4x4:
    1,275 dofs), or down to the thinnest cross-section without that stop"""
    from solidfem_amd import capi, mesh
    assert mesh.DIRECT_DOFS == capi.MG_DENSE_MAX
    m = BeamMesh(128, 32, 32)
    lv = beam_hierarchy(m)
    assert [l[1].shape[0] for l in lv] == [98304, 12288, 1536]
    assert [l[0].shape[0] for l in lv] == [65 * 17 * 17, 33 * 9 * 9, 17 * 5 * 5]
    assert 3 * lv[-1][0].shape[0] <= capi.MG_DENSE_MAX < 3 * lv[-2][0].shape[0]
    lv = beam_hierarchy(m, direct_dofs=None)
    assert [l[1].shape[0] for l in lv] == [98304, 12288, 1536, 192]
    assert [l[0].shape[0] for l in lv] == [65 * 17 * 17, 33 * 9 * 9, 17 * 5 * 5, 9 * 3 * 3]


def test_prolongation_nesting_and_orientation():
    m = BeamMesh(16, 8, 4, jitter=0.1, seed=7)
    lv = beam_hierarchy(m)
    assert len(lv) == 1
    X0, tets, fix, inj, ptr, col, val = lv[0]
    c = BeamMesh(8, 4, 2)
    P = dense_p(m.nn, X0.shape[0], ptr, col, val)
    # coarse nodes coincide with fine nodes, where P is the identity row
    assert np.array_equal(X0, m.X[inj])
    assert np.allclose(P[inj], np.eye(X0.shape[0]))
    assert np.allclose(P.sum(axis=1), 1.0)
    # exact on fields trilinear in the grid indices
    gf, gc = m.grid_index.astype(float), 2.0 * c.grid_index.astype(float)
    for f in (lambda g: 1.0 + 0 * g[:, 0], lambda g: 2 * g[:, 0] - g[:, 1] + 0.5 * g[:, 2],
              lambda g: g[:, 0] * g[:, 1] * g[:, 2]):
        assert np.abs(P @ f(gc) - f(gf)).max() < 1e-12
    # fixed flags are the fine flags at the coinciding nodes; positive volumes
    assert np.array_equal(fix, m.fixed_mask()[inj])
    Pt = X0[tets]
    vol = np.einsum("ij,ij->i", np.cross(Pt[:, 1] - Pt[:, 0], Pt[:, 2] - Pt[:, 0]), Pt[:, 3] - Pt[:, 0])
    assert np.all(vol > 0)
    # every coarse element lies in the fine mesh's box
    assert np.all(X0.min(axis=0) >= m.X.min(axis=0)) and np.all(X0.max(axis=0) <= m.X.max(axis=0))


def test_no_level_for_odd_or_thin_grids():
    assert beam_hierarchy(BeamMesh(10, 3, 3)) == []
    assert beam_hierarchy(BeamMesh(16, 2, 2)) == []


def test_coarse_moduli_are_cell_means():
    """per-element moduli of a coarse level: the mean over the fine tets whose
    centroids lie in each coarse hex (brute force over centroids)"""
    m = BeamMesh(8, 4, 4)
    rng = np.random.default_rng(3)
    E = 10 ** (5 + 2 * rng.random(m.ne))
    nu = 0.3 + 0.18 * rng.random(m.ne)
    lv = beam_hierarchy(m, young_e=E, poisson_e=nu)
    X0, tets, fx, inj, ptr, col, val, Ec, nuc = lv[0]
    c = BeamMesh(4, 2, 2)
    cf = m.X[m.tets].mean(axis=1)                 # fine centroids
    cc = X0[tets].mean(axis=1)                    # coarse centroids
    H = np.array(c.h)
    fcell = np.floor(cf / H).astype(int)
    ccell = np.floor(cc / H).astype(int)
    for t in range(c.ne):
        inside = np.all(fcell == ccell[t], axis=1)
        assert inside.sum() == 48
        assert np.isclose(Ec[t], E[inside].mean()) and np.isclose(nuc[t], nu[inside].mean())
    assert np.allclose(coarse_cell_mean(m, np.full(m.ne, 2.5)), 2.5)
    with pytest.raises(ValueError):
        beam_hierarchy(m, young_e=E)

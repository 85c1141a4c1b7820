"""Pins the FP64 CPU reference of the learned preconditioner (oracle/gnn_ref.py)
by central differences of its own loss, and checks its block construction:
the untrained network (zero last layer) is block Jacobi."""
import numpy as np

from oracle import gnn_ref as G
from oracle import oracle as O
from solidfem_amd.mesh import BeamMesh


def _problem():
    m = BeamMesh(4, 2, 2, 0.4, 0.2, 0.2)
    fixed = m.fixed_mask()
    p = O.OracleProblem(m.X, m.tets, fixed, O.NEO_HOOKEAN, 1e5, 0.4, 1000.0)
    B = p.block_diag(m.X)
    ptr, idx = G.node_graph(m.tets, m.nn)
    feat, cb = G.features(B, fixed, ptr, idx)
    return m, fixed, p, B, ptr, idx, feat, cb


def _params(seed, zero_last=False):
    rng = np.random.default_rng(seed)
    parts = []
    for li, (K, NO) in enumerate(G.LAYERS):
        scale = 0.0 if (zero_last and li == len(G.LAYERS) - 1) else np.sqrt(2.0 / K)
        parts.append(rng.standard_normal(K * NO) * scale)
        parts.append(rng.standard_normal(NO) * 0.05)
    if zero_last:
        parts[-1][:] = 0
    return np.concatenate(parts)


def test_untrained_network_is_block_jacobi():
    m, fixed, p, B, ptr, idx, feat, cb = _problem()
    net = G.GnnRef(_params(0, zero_last=True), ptr, idx, use_bf16=False)
    c = net.forward(feat, cb)
    Mi = net.blocks(c["O"], cb)
    free = fixed == 0
    J = np.linalg.inv(B[free])
    assert np.abs(Mi[free] - (1 + G.EPS) * J).max() < 1e-12 * np.abs(J).max()
    assert np.abs(Mi[~free]).max() == 0


def test_graph_and_features():
    m, fixed, p, B, ptr, idx, feat, cb = _problem()
    # symmetric graph, no self loops
    pairs = set()
    for i in range(m.nn):
        for j in idx[ptr[i]:ptr[i + 1]]:
            pairs.add((i, j))
            assert i != j
    assert all((j, i) in pairs for i, j in pairs)
    # mirrored 6-tet hexes: a node where eight hex diagonals meet has 18 neighbours
    assert 4 <= np.diff(ptr).min() and np.diff(ptr).max() <= 18
    # C C^T = D^-1 on free nodes
    C = G.lower(cb)
    free = fixed == 0
    D = 0.5 * (B + np.transpose(B, (0, 2, 1)))
    assert np.abs(C[free] @ np.transpose(C[free], (0, 2, 1)) - np.linalg.inv(D[free])).max() < 1e-10 * np.abs(np.linalg.inv(D[free])).max()


def test_reference_gradient_matches_finite_differences():
    m, fixed, p, B, ptr, idx, feat, cb = _problem()
    rng = np.random.default_rng(3)
    free = (fixed == 0)[:, None]
    R = rng.standard_normal((3, m.nn, 3)) * free
    # e = K^-1 r via CG to machine precision on this small problem
    er = np.array([np.sum(R[s] * p.cg(m.X, R[s], 5000, 1e-7)[0]) for s in range(3)])
    K = lambda z: p.matvec(m.X, z)
    w = _params(7) * 0.5
    net = G.GnnRef(w, ptr, idx, use_bf16=False)
    l0, g = net.loss(feat, cb, R, er, K)
    assert 0 < l0 < 1
    h = 1e-6
    for k in rng.choice(w.size, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        lp = G.GnnRef(wp, ptr, idx, use_bf16=False).loss(feat, cb, R, er, K, grad=False)[0]
        lm = G.GnnRef(wm, ptr, idx, use_bf16=False).loss(feat, cb, R, er, K, grad=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-3], dtype=np.float32)
    y = G.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0            # 1 + 2^-8 ties to even
    assert y[2] == 1.0078125                       # 1 + 1.5 * 2^-8 rounds up
    assert abs(y[3] + 3.140625) < 1e-12
    assert abs(y[4] - 1e-3) / 1e-3 < 2 ** -8

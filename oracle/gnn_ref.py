"""FP64 CPU reference of the learned block preconditioner (the network, its
block construction, the training loss and the parameter gradient).

TEST INFRASTRUCTURE ONLY: the FP64 reference for the device kernels of
solidfem_amd/csrc/sfem_gnn.hip and sfem_gnn_host.cpp.  The reference has no
counterpart (SolidFEM's only preconditioner is a probed scalar diagonal,
Math/IterativeSolvers.h:133-152); this restates the network defined in
include/sfem_capi.h.  Pinned by central differences of its own loss
(test_gnn_cpu.py); it then pins the device gradient.

With bf16=True, every GEMM operand (activations, weights, upstream gradients)
is rounded to bfloat16 (round to nearest even) exactly where the device rounds
it, and the products are accumulated in FP64.
"""
from __future__ import annotations

import numpy as np

WIDTH, FIN, FOUT, ROUNDS = 32, 16, 16, 3
EPS = 0.01
LAYERS = [(FIN, WIDTH), (WIDTH, WIDTH)] + [(2 * WIDTH, WIDTH), (WIDTH, WIDTH)] * ROUNDS + [(WIDTH, WIDTH), (WIDTH, FOUT)]


def bf16(x):
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    out, off = [], 0
    for K, NO in LAYERS:
        W = np.asarray(params[off:off + K * NO], dtype=np.float64).reshape(K, NO)
        off += K * NO
        b = np.asarray(params[off:off + NO], dtype=np.float64)
        off += NO
        out.append((W, b))
    assert off == len(params)
    return out


def pack(grads):
    return np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])


def node_graph(tets, nn):
    """CSR of nodes sharing an element, ascending, no self loops."""
    t = np.asarray(tets, dtype=np.int64)
    a = np.repeat(t, 4, axis=1).reshape(-1)
    b = np.tile(t, (1, 4)).reshape(-1)
    keep = a != b
    key = np.unique(a[keep] * nn + b[keep])
    src, dst = key // nn, key % nn
    ptr = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=nn))])
    return ptr, dst


def features(B, fixed, ptr, idx):
    """Per-node inputs and Cholesky factors of the inverse diagonal blocks."""
    N = B.shape[0]
    D = 0.5 * (B + np.transpose(B, (0, 2, 1)))
    t = np.trace(D, axis1=1, axis2=2) / 3.0
    t = np.where(t > 0, t, 1.0)
    f = np.zeros((N, FIN))
    f[:, 0], f[:, 1], f[:, 2] = D[:, 0, 0] / t, D[:, 1, 1] / t, D[:, 2, 2] / t
    f[:, 3], f[:, 4], f[:, 5] = D[:, 0, 1] / t, D[:, 0, 2] / t, D[:, 1, 2] / t
    deg = np.diff(ptr)
    lt = np.log(t)
    nbr = np.repeat(np.arange(N), deg)
    ln = np.bincount(nbr, weights=lt[idx], minlength=N)
    nf = np.bincount(nbr, weights=fixed[idx].astype(np.float64), minlength=N)
    with np.errstate(invalid="ignore", divide="ignore"):
        f[:, 6] = np.where(deg > 0, lt - ln / np.maximum(deg, 1), 0.0)
        f[:, 9] = np.where(deg > 0, nf / np.maximum(deg, 1), 0.0)
    f[:, 7] = deg / 16.0
    f[:, 8] = fixed.astype(np.float64)
    cb = np.zeros((N, 6))
    free = fixed == 0
    Di = np.linalg.inv(D[free])
    Lc = np.linalg.cholesky(Di)
    cb[free] = np.stack([Lc[:, 0, 0], Lc[:, 1, 1], Lc[:, 2, 2], Lc[:, 1, 0], Lc[:, 2, 0], Lc[:, 2, 1]], axis=1)
    return f, cb


def lower(v):
    L = np.zeros(v.shape[:-1] + (3, 3))
    L[..., 0, 0], L[..., 1, 1], L[..., 2, 2] = v[..., 0], v[..., 1], v[..., 2]
    L[..., 1, 0], L[..., 2, 0], L[..., 2, 1] = v[..., 3], v[..., 4], v[..., 5]
    return L


class GnnRef:
    def __init__(self, params, ptr, idx, use_bf16=True):
        self.P = unpack(params)
        self.ptr, self.idx = np.asarray(ptr), np.asarray(idx)
        self.deg = np.diff(self.ptr)
        self.rows = np.repeat(np.arange(len(self.deg)), self.deg)
        self.q = bf16 if use_bf16 else (lambda x: np.asarray(x, dtype=np.float64))

    def mean(self, h):
        N = h.shape[0]
        s = np.zeros_like(h)
        np.add.at(s, self.rows, h[self.idx])
        return np.where(self.deg[:, None] > 0, s / np.maximum(self.deg, 1)[:, None], 0.0)

    def mean_T(self, dm):
        w = dm / np.maximum(self.deg, 1)[:, None]
        s = np.zeros_like(dm)
        np.add.at(s, self.idx, w[self.rows])
        return s

    def gemm(self, X, l):
        W, b = self.P[l]
        return self.q(X) @ self.q(W) + b

    def forward(self, feat, cbase):
        c = {"feat": feat}
        c["A1"] = self.gemm(feat, 0)
        c["H1"] = np.maximum(c["A1"], 0)
        h = [self.gemm(c["H1"], 1)]
        c["M"], c["B"], c["C"] = [], [], []
        for r in range(ROUNDS):
            m = self.mean(h[r])
            U = np.concatenate([h[r], m], axis=1)
            Bv = self.gemm(U, 2 + 2 * r)
            Cv = np.maximum(Bv, 0)
            h.append(h[r] + self.gemm(Cv, 3 + 2 * r))
            c["M"].append(m)
            c["B"].append(Bv)
            c["C"].append(Cv)
        c["h"] = h
        c["A3"] = self.gemm(h[ROUNDS], 2 + 2 * ROUNDS)
        c["H3"] = np.maximum(c["A3"], 0)
        c["O"] = self.gemm(c["H3"], 3 + 2 * ROUNDS)
        c["cbase"] = cbase
        return c

    @staticmethod
    def blocks(O, cbase):
        ex = np.exp(np.clip(O[:, :3], -8, 8))
        L = lower(np.concatenate([ex, O[:, 3:6]], axis=1))
        C = lower(cbase)
        S = L @ np.transpose(L, (0, 2, 1)) + EPS * np.eye(3)
        return C @ S @ np.transpose(C, (0, 2, 1))

    def blocks_bwd(self, O, cbase, dM):
        ex = np.exp(np.clip(O[:, :3], -8, 8))
        L = lower(np.concatenate([ex, O[:, 3:6]], axis=1))
        C = lower(cbase)
        dS = np.transpose(C, (0, 2, 1)) @ dM @ C
        dL = (dS + np.transpose(dS, (0, 2, 1))) @ L
        dO = np.zeros_like(O)
        clamp = (O[:, :3] < -8) | (O[:, :3] > 8)
        dO[:, 0] = np.where(clamp[:, 0], 0, dL[:, 0, 0] * ex[:, 0])
        dO[:, 1] = np.where(clamp[:, 1], 0, dL[:, 1, 1] * ex[:, 1])
        dO[:, 2] = np.where(clamp[:, 2], 0, dL[:, 2, 2] * ex[:, 2])
        dO[:, 3], dO[:, 4], dO[:, 5] = dL[:, 1, 0], dL[:, 2, 0], dL[:, 2, 1]
        return dO

    def lin_bwd(self, X, dY, l):
        """gradients of Y = X W + b: (dW, db, dX) with the device's bf16 roundings"""
        W, _ = self.P[l]
        dW = self.q(X).T @ self.q(dY)
        db = dY.sum(axis=0)
        dX = self.q(dY) @ self.q(W).T
        return dW, db, dX

    def loss(self, feat, cbase, R, er, Kmul, grad=True):
        """R: (S, N, 3) residuals, er: (S,) r.e, Kmul(z) -> K z (N, 3)."""
        c = self.forward(feat, cbase)
        Mi = self.blocks(c["O"], cbase)
        S = R.shape[0]
        total = 0.0
        dM = np.zeros_like(Mi)
        for s in range(S):
            z = np.einsum("nij,nj->ni", Mi, R[s])
            q = Kmul(z)
            a, b = np.sum(z * R[s]), np.sum(z * q)
            total += 1.0 - a * a / (b * er[s])
            f = 2.0 * a / (b * er[s]) / S
            g = f * (a / b) * q - f * R[s]
            dM += np.einsum("ni,nj->nij", g, R[s])
        total /= S
        if not grad:
            return total, None
        G = [None] * len(LAYERS)
        dO = self.blocks_bwd(c["O"], cbase, dM)
        L = len(LAYERS)
        dW, db, dH3 = self.lin_bwd(c["H3"], dO, L - 1)
        G[L - 1] = (dW, db)
        dA3 = np.where(c["A3"] > 0, dH3, 0.0)
        dW, db, dh = self.lin_bwd(c["h"][ROUNDS], dA3, L - 2)
        G[L - 2] = (dW, db)
        for r in range(ROUNDS - 1, -1, -1):
            dW, db, dC = self.lin_bwd(c["C"][r], dh, 3 + 2 * r)
            G[3 + 2 * r] = (dW, db)
            dB = np.where(c["B"][r] > 0, dC, 0.0)
            U = np.concatenate([c["h"][r], c["M"][r]], axis=1)
            dW, db, dU = self.lin_bwd(U, dB, 2 + 2 * r)
            G[2 + 2 * r] = (dW, db)
            dh = dh + dU[:, :WIDTH] + self.mean_T(dU[:, WIDTH:])
        dW, db, dH1 = self.lin_bwd(c["H1"], dh, 1)
        G[1] = (dW, db)
        dA1 = np.where(c["A1"] > 0, dH1, 0.0)
        dW, db, _ = self.lin_bwd(feat, dA1, 0)
        G[0] = (dW, db)
        return total, pack(G)

"""Structured tetrahedral beam meshes and the loads the beam workloads use.

The hex -> 6 tet split and the node numbering follow the reference's own beam
fixture ``Models/box2.tet`` (an 8x2x2 grid) as it comes out of
``IO::LoadFromTet1File`` (Source/FemIO.cpp:506-581, which swaps file y/z on
read, :548):

* node id ``n = ix*(nz+1)*(ny+1) + iz*(ny+1) + iy``;
* hexes are visited ``for iy: for iz: for ix`` (x fastest);
* each hex is cut into 6 tets around one body diagonal.  The diagonal's start
  corner has offsets ``(ix%2, iy%2, 1 - iz%2)``, so neighbouring hexes mirror
  and faces match.  The six tets are (D, ring[k], ring[k+1], D') over the ring
  of the remaining six corners.

Vertex order inside each tet is chosen so the signed volume is positive (the
reference repairs orientation in FemBody::Prepare, Source/FemBody.cpp:127-156).
"""
from __future__ import annotations

import numpy as np

# ring of the six corners around the 000-111 diagonal, as offsets XOR'ed with D
_RING = np.array([[1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1], [1, 0, 1]], dtype=np.int64)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """Counter-based hash (splitmix64 finaliser) -> uint64."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n deterministic uniforms in [0,1) from (seed, stream, index)."""
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = idx + np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    h = _splitmix64(key)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class BeamMesh:
    """Hex grid nx*ny*nz over [0,Lx]x[0,Ly]x[0,Lz], 6 tets per hex."""

    def __init__(self, nx, ny, nz, Lx=1.0, Ly=0.25, Lz=0.25, jitter=0.0, seed=0):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.L = (float(Lx), float(Ly), float(Lz))
        hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
        self.h = (hx, hy, hz)
        NX, NY, NZ = nx + 1, ny + 1, nz + 1
        ix, iz, iy = np.meshgrid(np.arange(NX), np.arange(NZ), np.arange(NY), indexing="ij")
        # node id = ix*NZ*NY + iz*NY + iy  <=>  C-order over (ix, iz, iy)
        X = np.stack([ix * hx, iy * hy, iz * hz], axis=-1).reshape(-1, 3).astype(np.float64)
        self.grid_index = np.stack([ix, iy, iz], axis=-1).reshape(-1, 3)
        if jitter > 0:
            X = self._jitter(X, jitter, seed)
        self.X = X
        self.tets = self._tets()
        self.nn = X.shape[0]
        self.ne = self.tets.shape[0]

    def node_id(self, ix, iy, iz):
        return (ix * (self.nz + 1) + iz) * (self.ny + 1) + iy

    def _jitter(self, X, amp, seed):
        """Move nodes by U(-amp, amp) cells per axis; nodes on a boundary face
        move only within that face."""
        g = self.grid_index
        n = X.shape[0]
        out = X.copy()
        dims = (self.nx, self.ny, self.nz)
        for a in range(3):
            u = uniform01(seed, 11 + a, n) * 2.0 - 1.0
            on_face = (g[:, a] == 0) | (g[:, a] == dims[a])
            out[:, a] += np.where(on_face, 0.0, u * amp * self.h[a])
        return out

    def _tets(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        iy, iz, ix = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
        ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()          # hex order: iy, iz, ix(fastest)
        D = np.stack([ix % 2, iy % 2, 1 - iz % 2], axis=1)
        Dbar = 1 - D
        corners = []
        for k in range(6):
            a = _RING[k] ^ D
            b = _RING[(k + 1) % 6] ^ D
            corners.append((D, a, b, Dbar))
        out = np.empty((ix.size, 6, 4), dtype=np.int64)
        for k, quad in enumerate(corners):
            for v, off in enumerate(quad):
                out[:, k, v] = self.node_id(ix + off[:, 0], iy + off[:, 1], iz + off[:, 2])
        tets = out.reshape(-1, 4)
        # orientation: positive signed volume
        P = self.X[tets]
        vol = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
        neg = vol < 0
        tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
        return np.ascontiguousarray(tets.astype(np.int32))

    # ------------------------------------------------------------ boundary sets
    def fixed_mask(self):
        """Nodes on the x = 0 face (the clamped end)."""
        return (self.grid_index[:, 0] == 0).astype(np.uint8)

    def tip_triangles(self):
        """Tet faces lying on x = Lx, oriented with outward (+x) normals."""
        on = self.grid_index[:, 0] == self.nx
        t = self.tets
        faces = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
        # keep hex order: sort by source tet index
        tri_all = []
        for f in faces:
            tri = t[:, f]
            sel = np.nonzero(on[tri].all(axis=1))[0]
            tri_all.append(np.stack([sel, tri[sel][:, 0], tri[sel][:, 1], tri[sel][:, 2]], axis=1))
        T = np.concatenate(tri_all, axis=0)
        T = T[np.argsort(T[:, 0], kind="stable")][:, 1:]
        P = self.X[T]
        n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        flip = n[:, 0] < 0
        T[flip, 1], T[flip, 2] = T[flip, 2].copy(), T[flip, 1].copy()
        return np.ascontiguousarray(T.astype(np.int32))


def traction_nodal_forces(X, tris, traction_fn):
    """Consistent nodal loads of a dead traction t(p) that is linear over each
    triangle: f_i = A/12 (2 t_i + t_j + t_k).  Returns an (nn,3) array."""
    f = np.zeros_like(X)
    P = X[tris]
    area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    T = np.stack([traction_fn(P[:, k]) for k in range(3)], axis=1)   # (nt,3,3)
    for k in range(3):
        fk = (area / 12.0)[:, None] * (2 * T[:, k] + T[:, (k + 1) % 3] + T[:, (k + 2) % 3])
        np.add.at(f, tris[:, k], fk)
    return f


def torsion_traction(center, tau):
    """Tangential dead traction t(p) = tau * e_x x (p - c) on an x = const face."""
    c = np.asarray(center, dtype=np.float64)

    def fn(p):
        r = p - c
        return tau * np.stack([np.zeros(len(p)), -r[:, 2], r[:, 1]], axis=1)
    return fn


def lame(E, nu):
    """mu = E/(2(1+nu)), lambda = E nu/((1+nu)(1-2nu))  (Source/FemPhysicsBase.cpp:40-41)."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = 0.5 * E / (1.0 + nu)
    return mu, lam


def load_tet1(path, scale=1.0):
    """Read a `.tet` file the way IO::LoadFromTet1File does (Source/FemIO.cpp:506-581):
    'v x y z' lines (y and z swapped on read) and 't a b c d' lines."""
    X, T = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("v"):
                x, y, z = (float(v) for v in line.split()[1:4])
                X.append((x * scale, z * scale, y * scale))
            elif line.startswith("t"):
                T.append(tuple(int(v) for v in line.split()[1:5]))
    return np.array(X, dtype=np.float64), np.array(T, dtype=np.int32)


def coarse_cell_mean(fine, values):
    """per-element values of `fine` (a BeamMesh, 6 tets per hex in hex order
    iy, iz, ix) -> per-element values of the beam with half the cells in every
    direction: the mean over the 48 fine tets inside each coarse hex, given
    to its 6 tets"""
    v = np.asarray(values, dtype=np.float64).reshape(fine.ny, fine.nz, fine.nx, 6).mean(axis=3)
    v = v.reshape(fine.ny // 2, 2, fine.nz // 2, 2, fine.nx // 2, 2).mean(axis=(1, 3, 5))
    return np.repeat(v.reshape(-1), 6)


# the library solves the coarsest multigrid level directly when it has at
# most this many dofs (3 x nodes; MG_DENSE_MAX of the library, capi.MG_DENSE_MAX)
DIRECT_DOFS = 1536


def beam_hierarchy(m, min_cells=2, fixed=None, young_e=None, poisson_e=None, direct_dofs=DIRECT_DOFS):
    """Nested coarse levels of a beam mesh for the multigrid preconditioner:
    each level halves the hex grid in every direction while all three cell
    counts stay even and the thinner cross-section keeps >= min_cells cells,
    and stops at the first level of at most direct_dofs dofs (3 x nodes),
    which the library solves directly (None: no such stop).
    Coarse node (i, j, k) sits at fine node (2i, 2j, 2k) (its position, so a
    jittered mesh stays nested), and the prolongation interpolates node
    values trilinearly in the grid indices.  Per-element moduli (young_e,
    poisson_e) are averaged over each coarse hex (coarse_cell_mean).
    Returns [(X0, tets, fixed, inject, p_ptr, p_col, p_val[, young_e,
    poisson_e])], finest coarse level first (the arguments of
    capi.Solver.add_mg_level)."""
    levels = []
    fine, ffix = m, (m.fixed_mask() if fixed is None else np.asarray(fixed, dtype=np.uint8))
    if (young_e is None) != (poisson_e is None):
        raise ValueError("beam_hierarchy: give both young_e and poisson_e, or neither")
    fE = None if young_e is None else np.asarray(young_e, dtype=np.float64)
    fnu = None if poisson_e is None else np.asarray(poisson_e, dtype=np.float64)
    while (fine.nx % 2 == 0 and fine.ny % 2 == 0 and fine.nz % 2 == 0
           and min(fine.ny, fine.nz) // 2 >= min_cells):
        c = BeamMesh(fine.nx // 2, fine.ny // 2, fine.nz // 2, *fine.L)
        g = c.grid_index
        inject = fine.node_id(2 * g[:, 0], 2 * g[:, 1], 2 * g[:, 2]).astype(np.int32)
        c.X = fine.X[inject].copy()
        tets = c.tets.astype(np.int64)
        P = c.X[tets]
        vol = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
        neg = vol < 0
        tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
        c.tets = np.ascontiguousarray(tets.astype(np.int32))
        cfix = ffix[inject].astype(np.uint8)
        # trilinear prolongation rows (fine nodes) over the coarse grid
        gf = fine.grid_index
        lo = gf // 2
        odd = gf % 2
        rows, cols, vals = [], [], []
        for cx in (0, 1):
            for cy in (0, 1):
                for cz in (0, 1):
                    sel = np.array([cx, cy, cz])
                    ok = np.all((sel == 0) | (odd == 1), axis=1)
                    w = np.prod(np.where(odd == 1, 0.5, 1.0), axis=1)
                    idx = lo + sel * odd
                    n = np.nonzero(ok)[0]
                    rows.append(n)
                    cols.append(c.node_id(idx[n, 0], idx[n, 1], idx[n, 2]))
                    vals.append(w[n])
        rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
        order = np.lexsort((cols, rows))
        rows, cols, vals = rows[order], cols[order], vals[order]
        p_ptr = np.zeros(fine.nn + 1, dtype=np.int64)
        np.add.at(p_ptr, rows + 1, 1)
        p_ptr = np.cumsum(p_ptr)
        lv = (c.X, c.tets, cfix, inject, p_ptr, cols.astype(np.int32), vals.astype(np.float64))
        if fE is not None:
            fE, fnu = coarse_cell_mean(fine, fE), coarse_cell_mean(fine, fnu)
            lv = lv + (fE, fnu)
        levels.append(lv)
        fine, ffix = c, cfix
        if direct_dofs is not None and 3 * c.nn <= direct_dofs:
            break
    return levels


def _tet_volumes(a, b, c, d):
    """signed volumes det[b - a, c - a, d - a] / 6 (rows of points)"""
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6.0


def barycentric(p, X, tets):
    """(ne, 4) barycentric weights of point p in every tetrahedron: sub-volume
    ratios as ComputeBarycentric computes them (Source/FemBody.cpp:17-29)"""
    x0, x1, x2, x3 = (X[tets[:, k]] for k in range(4))
    P = np.broadcast_to(p, x0.shape)
    vol = _tet_volumes(x0, x1, x2, x3)
    return np.stack([_tet_volumes(P, x1, x2, x3), _tet_volumes(x0, P, x2, x3), _tet_volumes(x0, x1, P, x3),
                     _tet_volumes(x0, x1, x2, P)], axis=1) / vol[:, None]


def create_cable(X, tets, divs=10, length=1.0, offset=(0.0, 0.0, 0.0), direction=(1.0, 0.0, 0.0), stiffness=1000.0,
                 use_free_cable=True, scale=1.0):
    """CreateCable (Source/FemBody.cpp:31-104) restated: divs + 1 points
    offset + i (length / divs) direction, each registered to the element with
    the smallest score sum_j |100 X_j - p|^2 among those whose barycentric
    weights all exceed -0.001 (else among all elements), at its barycentric
    weights.  Returns dict(elem, bary (n, 3), rest_length, stiffness, free)
    where `free` marks points further outside their element than 0.1 in a
    weight (the reference's free cable nodes; sfem_add_cable refuses them)."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    tets = np.asarray(tets).reshape(-1, 4)
    div = length / divs
    d = np.asarray(direction, dtype=np.float64)
    off = np.asarray(offset, dtype=np.float64)
    elem, bary, free = [], [], []
    Xs = scale * X
    for i in range(divs + 1):
        p = i * div * d + off
        w = barycentric(p, Xs, tets)
        score = np.sum((100.0 * X[tets] - p) ** 2, axis=(1, 2))
        e = int(np.argmin(score))
        positive = np.all(w > -0.001, axis=1)
        if positive.any():
            e = int(np.flatnonzero(positive)[np.argmin(score[positive])])
        we = w[e]
        elem.append(e)
        bary.append(we[:3])
        free.append(bool(use_free_cable and np.any(we < -0.1)))
    return {"elem": np.array(elem, dtype=np.int32), "bary": np.array(bary), "rest_length": div / scale,
            "stiffness": float(stiffness), "free": np.array(free)}

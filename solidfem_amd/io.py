"""On-disk formats of SolidFEM scenes, for driving the device solver from the
reference's own files:

- `.tet` meshes (IO::LoadFromTet1File, Source/FemIO.cpp:506-581),
- `<solidfem>` XML scenes (IO::LoadFromXmlFile, Source/FemIO.cpp:583-847),
- legacy-VTK export with the displacement / stress heat map
  (IO::ExportToVTKHeatMap and IO::ExportToVTKFile, Source/FemIO.cpp:317-399).

Only what the matrix-free path uses is read: `.vol` / `.feb` models and cables
raise a clear error rather than being ignored.
"""
from __future__ import annotations

import os
import re
import xml.etree.ElementTree as ET

import numpy as np

from .mesh import load_tet1

# MaterialModelType / SimulationType values (Include/FemDataStructures.h:71-88)
_MATERIALS = {"linear": 0, "stvk": 1, "corotational": 2, "neo-hookean": 3}
_SIMTYPES = {"static": 0, "quasi-static": 1, "explicit": 2, "implicit": 3}


def _ints(text):
    """the reference's std::stoi loop: integers separated by blanks or commas"""
    return [int(t) for t in re.findall(r"-?\d+", text or "")]


def load_xml(path):
    """Parse a `<solidfem>` scene.  Returns dict(nodes (nn, 3), tets (ne, 4),
    fixed (node ids), surf_tris (flat ids), config) where `config` uses the
    FemConfig fields the reference sets: material, young, poisson, density,
    simtype, force_step or substeps, gravity, pressure, outer_iterations,
    abs_tolerance, scale, and the pysolidfem-style keys derived from them."""
    root = ET.parse(path).getroot()
    if root.tag != "solidfem":
        raise ValueError(f"{path}: not a <solidfem> scene")
    base = os.path.dirname(os.path.abspath(path))
    cfg = {}
    nodes = tets = None
    fixed, surf = [], []
    model = root.find("model")
    if model is not None:
        mpath = model.get("path")
        if not mpath:
            raise ValueError(f"{path}: <model> without a path")
        scale = float(model.get("scale", "1"))
        kind = model.get("type")
        if kind != "tet1":
            raise ValueError(f"{path}: model type {kind!r} is not supported here (tet1 only)")
        nodes, tets = load_tet1(os.path.join(base, mpath), scale)
        swap = model.get("swap")
        pairs = {"xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}
        if swap in pairs:
            a, b = pairs[swap]
            nodes[:, [a, b]] = nodes[:, [b, a]]
        d = model.find("dirichlet")
        if d is not None:
            fixed = _ints(d.text)
        cfg["scale"] = scale
    mat = root.find("material")
    if mat is not None:
        if mat.get("name") in _MATERIALS:
            cfg["material"] = _MATERIALS[mat.get("name")]
        for key, attr in (("young", "young"), ("poisson", "poisson"), ("density", "density")):
            if mat.get(attr) is not None:
                cfg[key] = float(mat.get(attr))
    sim = root.find("simulation")
    if sim is not None:
        st = sim.get("type")
        if st in _SIMTYPES:
            cfg["simtype"] = _SIMTYPES[st]
        if sim.get("method") == "mixed":
            cfg["mixed"] = True
        if sim.get("steps") is not None:
            n = int(sim.get("steps"))
            if cfg.get("simtype", 0) < 2:
                cfg["force_step"] = 1.0 / n
            else:
                cfg["substeps"] = n
        if sim.get("gravity") is not None:
            cfg["gravity"] = float(sim.get("gravity"))
        pr = sim.find("pressure")
        if pr is not None:
            if pr.get("p") is not None:
                cfg["pressure"] = float(pr.get("p"))
            surf = _ints(pr.text)
    sol = root.find("solver")
    if sol is not None:
        if sol.get("iterations") is not None:
            cfg["maxiter"] = int(sol.get("iterations"))
        if sol.get("tolerance") is not None:
            cfg["tol"] = float(sol.get("tolerance"))
    if root.find("cable") is not None:
        raise ValueError(f"{path}: cables are not on the matrix-free path")
    return {"nodes": nodes, "tets": tets, "fixed": np.asarray(fixed, dtype=np.int64),
            "surf_tris": np.asarray(surf, dtype=np.int64), "config": cfg}


def prepare(nodes, tets):
    """FemBody::Prepare (Source/FemBody.cpp:127-156): swap the first two
    indices of every tetrahedron with negative signed volume, so that all
    volumes are positive (the y/z swap of `.tet` files mirrors meshes).
    Returns the new (ne, 4) index array."""
    X = np.asarray(nodes, dtype=np.float64).reshape(-1, 3)
    T = np.array(tets, dtype=np.int64).reshape(-1, 4)
    D = np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=2)
    neg = np.linalg.det(D) < 0
    T[neg, 0], T[neg, 1] = T[neg, 1].copy(), T[neg, 0].copy()
    return T


def export_vtk(path, X0, tets, x, stress=None):
    """Legacy ASCII VTK unstructured grid of the deformed mesh x (nn, 3):
    POINTS, CELLS (VTK_TETRA = 10), and with `stress` (ne, 3, 3) the heat map:
    POINT_DATA displacement vectors x - X0 and CELL_DATA stress tensors."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    tets = np.asarray(tets, dtype=np.int64).reshape(-1, 4)
    nn, ne = x.shape[0], tets.shape[0]
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nUnstructured Grid\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {nn} double\n")
        np.savetxt(f, x, fmt="%.17g")
        f.write(f"\nCELLS {ne} {5 * ne}\n")
        np.savetxt(f, np.hstack([np.full((ne, 1), 4), tets]), fmt="%d")
        f.write(f"\nCELL_TYPES {ne}\n")
        np.savetxt(f, np.full(ne, 10), fmt="%d")
        if stress is not None:
            u = x - np.asarray(X0, dtype=np.float64).reshape(-1, 3)
            f.write(f"\nPOINT_DATA {nn}\nVECTORS displacement float\n")
            np.savetxt(f, u, fmt="%.9g")
            f.write(f"\nCELL_DATA {ne}\nTENSORS stress float\n")
            np.savetxt(f, np.asarray(stress, dtype=np.float64).reshape(-1, 3), fmt="%.9g")


def read_vtk(path):
    """-> dict(points, cells, displacement?, stress?) of a file export_vtk wrote"""
    with open(path) as f:
        tok = f.read().split()
    out, i = {}, 0
    while i < len(tok):
        t = tok[i]
        if t == "POINTS":
            n = int(tok[i + 1])
            out["points"] = np.array(tok[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif t == "CELLS":
            n = int(tok[i + 1])
            out["cells"] = np.array(tok[i + 3:i + 3 + 5 * n], dtype=np.int64).reshape(n, 5)[:, 1:]
            i += 3 + 5 * n
        elif t == "VECTORS":
            n = out["points"].shape[0]
            out["displacement"] = np.array(tok[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif t == "TENSORS":
            n = out["cells"].shape[0]
            out["stress"] = np.array(tok[i + 3:i + 3 + 9 * n], dtype=np.float64).reshape(n, 3, 3)
            i += 3 + 9 * n
        else:
            i += 1
    return out

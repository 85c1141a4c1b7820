"""ctypes front-end of the FP64 CPU oracle (sfem_oracle.c) and of the compiled
reference (oracle/_ref/libsfem_ref.so, present only where the reference
checkout was available at build time).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare against.  The product (solidfem_amd/)
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsfem_ref.so")

LINEAR, STVK, COROTATIONAL, NEO_HOOKEAN = 0, 1, 2, 3
COROTATIONAL_POLAR = 16     # no reference counterpart: the polar rotation, consistent tangent

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class _Problem(C.Structure):
    _fields_ = [
        ("ne", C.c_int64), ("nn", C.c_int64),
        ("tets", _ip), ("X0", _dp), ("fixed", C.POINTER(C.c_uint8)),
        ("DmInv", _dp), ("vol", _dp), ("mu", _dp), ("lam", _dp), ("mass", _dp),
        ("material", C.c_int32), ("nthreads", C.c_int32),
        ("grav", C.c_double * 3),
        ("ntri", C.c_int64), ("tris", _ip), ("pressure", C.c_double),
        ("fext", _dp),
        ("inc_ptr", C.POINTER(C.c_int64)), ("inc_idx", _ip),
    ]


def build(force=False):
    """Compile liboracle.so (and the reference library when REF is present)."""
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "sfem_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])


TEMPLATES_PATH = os.path.join(HERE, "_ref", "ref_templates")


def build_ref():
    """the compiled reference library and, when libsfem_hip.so is built, the
    driver that instantiates the reference's solver templates on this
    library's C++ class (oracle/ref_templates.cpp)"""
    ref_root = os.environ.get("SFEM_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref_root, "Source")):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref", f"REF={ref_root}"])
        if os.path.exists(os.path.join(os.path.dirname(HERE), "solidfem_amd", "libsfem_hip.so")):
            subprocess.check_call(["make", "-s", "-C", HERE, "templates", f"REF={ref_root}"])
    return os.path.exists(REF_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(_Problem)
        for name, args, res in [
            ("orc_geometry", [P, _dp, _dp], None),
            ("orc_lumped_mass", [P, C.c_double, _dp], None),
            ("orc_stress", [C.c_int, C.c_double, C.c_double, _dp, _dp], None),
            ("orc_dstress", [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp], None),
            ("orc_energy_density", [C.c_int, C.c_double, C.c_double, _dp], C.c_double),
            ("orc_internal_forces", [P, _dp, _dp], None),
            ("orc_energy", [P, _dp], C.c_double),
            ("orc_matvec", [P, _dp, _dp, _dp], None),
            ("orc_pressure_forces", [P, _dp, C.c_double, _dp], None),
            ("orc_external", [P, _dp, C.c_double, _dp], None),
            ("orc_gradients", [P, _dp, C.c_double, _dp], None),
            ("orc_block_diag", [P, _dp, _dp], None),
            ("orc_cg", [P, _dp, _dp, _dp, C.c_int, C.c_float, _dp], C.c_int),
            ("orc_newton_cg_ref", [P, _dp, C.c_double, C.c_int, C.c_int, _ip], None),
            ("orc_newton_solve", [P, _dp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, _dp], C.c_int),
            ("orc_cable_forces", [C.c_int64, _ip, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                  C.c_double, _dp], None),
            ("orc_cable_energy", [C.c_int64, _ip, _dp, C.c_double, C.c_double, C.c_double, _dp], C.c_double),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def explicit_steps(problem, x, v, h, nsub, gravity):
    """SubStep (FemPhysicsMatrixFree.cpp:371-387) restated: nsub symplectic Euler
    substeps v += h f_int / m + h g, x += h v on free nodes (elastic forces
    only), with the oracle's FP64 internal forces and lumped masses."""
    x = _f64(x).reshape(-1, 3).copy()
    v = _f64(v).reshape(-1, 3).copy()
    free = problem.fixed == 0
    g = np.asarray(gravity, dtype=np.float64)
    for _ in range(nsub):
        f = problem.internal_forces(x)
        v[free] += (h / problem.mass[free])[:, None] * f[free] + h * g
        x[free] += h * v[free]
    return x, v


class ImplicitStep:
    """Backward-Euler terms of FemPhysicsMatrixFree with mTimeStep = h
    restated around an OracleProblem: residual r(x) = r_static(x) + m/h v -
    m/h^2 (x - x0) and operator K d + m/h^2 d on free nodes
    (ComputeGradients / MatrixVectorMultiply, FemPhysicsMatrixFree.cpp:523-560),
    and a dense FP64 Newton solve of r(x) = 0 for small meshes."""

    def __init__(self, problem, x0, v, h, frac=1.0):
        self.p, self.h, self.frac = problem, float(h), frac
        self.x0 = _f64(x0).reshape(-1, 3).copy()
        self.v = _f64(v).reshape(-1, 3).copy()
        self.free = problem.fixed == 0
        self.m = problem.mass

    def residual(self, x):
        r = self.p.gradients(x, self.frac)
        f = self.free
        r[f] += (self.m[f] / self.h)[:, None] * self.v[f] - (self.m[f] / self.h ** 2)[:, None] * (x[f] - self.x0[f])
        return r

    def matvec(self, x, d):
        q = self.p.matvec(x, d)
        f = self.free
        q[f] += (self.m[f] / self.h ** 2)[:, None] * d[f]
        return q

    def solve(self, x, tol=1e-12, max_newton=30):
        """Newton with the operator assembled column by column (tiny meshes)"""
        x = _f64(x).reshape(-1, 3).copy()
        dofs = np.flatnonzero(np.repeat(self.free, 3))
        for it in range(max_newton):
            r = self.residual(x)
            if np.linalg.norm(r) <= tol * max(np.linalg.norm(self.residual(self.x0)), 1e-300):
                return x, it
            K = np.zeros((dofs.size, dofs.size))
            for j, dof in enumerate(dofs):
                e = np.zeros(x.size)
                e[dof] = 1.0
                K[:, j] = self.matvec(x, e.reshape(-1, 3)).reshape(-1)[dofs]
            dx = np.zeros(x.size)
            dx[dofs] = np.linalg.solve(K, r.reshape(-1)[dofs])
            x += dx.reshape(-1, 3)
        return x, max_newton


def stress(mat, mu, lam, F):
    F = _f64(F).reshape(9)
    P = np.zeros(9)
    lib().orc_stress(mat, mu, lam, _p(F), _p(P))
    return P.reshape(3, 3)


def dstress(mat, mu, lam, F, dF):
    F = _f64(F).reshape(9)
    dF = _f64(dF).reshape(9)
    dP = np.zeros(9)
    lib().orc_dstress(mat, mu, lam, _p(F), _p(dF), _p(dP))
    return dP.reshape(3, 3)


def energy_density(mat, mu, lam, F):
    return lib().orc_energy_density(mat, mu, lam, _p(_f64(F).reshape(9)))


class OracleCable:
    """One cable for the oracle: n nodes inside elements elem at barycentric
    weights bary (n, 3) of the mesh tets, springs of rest length `rest`
    (CreateCable's layout, FemBody.cpp:31-104; AddCable's fields,
    FemDataStructures.h:212-220)."""

    def __init__(self, tets, elem, bary, rest, stiffness=1000.0, damping=0.0, actuation=1.0):
        tets = np.asarray(tets).reshape(-1, 4)
        self.elem = np.ascontiguousarray(elem, dtype=np.int32)
        self.bary = _f64(bary).reshape(-1, 3)
        self.node4 = np.ascontiguousarray(tets[self.elem], dtype=np.int32)
        w3 = 1.0 - self.bary.sum(axis=1)
        self.w4 = np.ascontiguousarray(np.column_stack([self.bary, w3]))
        self.rest, self.stiffness, self.damping, self.actuation = float(rest), float(stiffness), float(damping), \
            float(actuation)

    def forces(self, x, frac=1.0, v=None):
        """frac * the cable's node forces on the mesh nodes (nn, 3)"""
        x = _f64(x).reshape(-1, 3)
        f = np.zeros_like(x)
        vv = None if v is None else _f64(v).reshape(-1, 3)
        lib().orc_cable_forces(len(self.elem), _p(self.node4, _ip), _p(self.w4), self.rest, self.actuation,
                               self.stiffness, self.damping, _p(x), _p(vv), float(frac), _p(f))
        return f

    def energy(self, x):
        x = _f64(x).reshape(-1, 3)
        return lib().orc_cable_energy(len(self.elem), _p(self.node4, _ip), _p(self.w4), self.rest, self.actuation,
                                      self.stiffness, _p(x))

    def positions(self, x):
        x = _f64(x).reshape(-1, 3)
        return np.einsum("nk,nkc->nc", self.w4, x[self.node4])


def reference_total_energy(problem, x, frac, gravity, tris=None, pressure=0.0, cables=(), h=0.0):
    """FemPhysicsMatrixFree::ComputeEnergy (FemPhysicsMatrixFree.cpp:882-916)
    restated term by term, with its conventions: the elastic energy; minus
    m frac g_y y over free nodes (the y components only, whatever the gravity
    axis); minus (0.5 / h^2) m |x - x0|^2 when h > 0; plus |det[p1, p2, p3]|
    times (float) 1/6 times the pressure for the triangles 0, 3, 6, ... of the traction surface
    (the loop steps by 3 over the triangle count); plus the cable springs'
    potential.  `gravity` is the gravity vector (FemPhysicsBase.cpp:30-33)."""
    x = _f64(x).reshape(-1, 3)
    val = problem.energy(x)
    free = problem.fixed == 0
    m = problem.mass[free]
    val -= np.sum(m * frac * gravity[1] * x[free, 1])
    if h:
        val -= np.sum((0.5 / h / h) * m * np.sum((x[free] - problem.X0[free]) ** 2, axis=1))
    if tris is not None and pressure:
        t = np.asarray(tris).reshape(-1, 3)
        for i in range(0, t.shape[0], 3):
            p1, p2, p3 = x[t[i, 0]], x[t[i, 1]], x[t[i, 2]]
            # (1.f / 6.f) is a float constant there
            val += abs(float(np.float32(1.0) / np.float32(6.0)) * np.dot(p3, np.cross(p1, p2))) * pressure
    for c in cables:
        val += c.energy(x)
    return float(val)


class OracleProblem:
    """One mesh + material + loads, in the caller's original node order."""

    def __init__(self, X0, tets, fixed, material, young=None, poisson=None, density=1070.0,
                 gravity=(0.0, 0.0, 0.0), tris=None, pressure=0.0, fext=None,
                 mu=None, lam=None, nthreads=0, incidence=False):
        self.X0 = _f64(X0).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        self.fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        self.nn, self.ne = self.X0.shape[0], self.tets.shape[0]
        if mu is None:
            E = np.broadcast_to(np.asarray(young, dtype=np.float64), (self.ne,))
            nu = np.broadcast_to(np.asarray(poisson, dtype=np.float64), (self.ne,))
            # FemPhysicsBase.cpp:40-41
            lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
            mu = 0.5 * E / (1.0 + nu)
        # own, writable copies (the C side reads them through these pointers)
        self.mu = np.array(np.broadcast_to(mu, (self.ne,)), dtype=np.float64)
        self.lam = np.array(np.broadcast_to(lam, (self.ne,)), dtype=np.float64)
        self.material = int(material)
        self.DmInv = np.zeros((self.ne, 9))
        self.vol = np.zeros(self.ne)
        self.mass = np.zeros(self.nn)
        self.tris = None if tris is None else np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        self.fext = None if fext is None else _f64(fext).reshape(-1, 3)
        s = _Problem()
        s.ne, s.nn = self.ne, self.nn
        s.tets = _p(self.tets, _ip)
        s.X0 = _p(self.X0)
        s.fixed = _p(self.fixed, C.POINTER(C.c_uint8))
        s.DmInv, s.vol, s.mu, s.lam, s.mass = (_p(a) for a in (self.DmInv, self.vol, self.mu, self.lam, self.mass))
        s.material = self.material
        s.nthreads = int(nthreads)
        s.grav[:] = list(gravity)
        s.ntri = 0 if self.tris is None else self.tris.shape[0]
        s.tris = _p(self.tris, _ip) if self.tris is not None else None
        s.pressure = float(pressure)
        s.fext = _p(self.fext) if self.fext is not None else None
        self.inc_ptr = self.inc_idx = None
        if incidence:
            slots = np.argsort(self.tets.reshape(-1), kind="stable")
            counts = np.bincount(self.tets.reshape(-1), minlength=self.nn)
            self.inc_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
            self.inc_idx = slots.astype(np.int32)
            s.inc_ptr = _p(self.inc_ptr, C.POINTER(C.c_int64))
            s.inc_idx = _p(self.inc_idx, _ip)
        self.s = s
        L = lib()
        L.orc_geometry(C.byref(s), s.DmInv, s.vol)
        L.orc_lumped_mass(C.byref(s), float(density), s.mass)

    def _x(self, x):
        return _f64(self.X0 if x is None else x).reshape(-1, 3)

    def internal_forces(self, x=None):
        x = self._x(x)
        f = np.zeros_like(x)
        lib().orc_internal_forces(C.byref(self.s), _p(x), _p(f))
        return f

    def energy(self, x=None):
        return lib().orc_energy(C.byref(self.s), _p(self._x(x)))

    def dead_work(self, x=None, frac=1.0):
        """frac sum over free nodes of (m g + f_dead) . u, u = x - X0"""
        u = self._x(x) - self.X0
        free = self.fixed == 0
        f = self.mass[:, None] * np.asarray(self.s.grav[:])[None, :]
        if self.fext is not None:
            f = f + self.fext
        return float(frac * np.sum(f[free] * u[free]))

    def force_param_grads(self, x=None):
        """d f_int / d mu and d f_int / d lambda (nn, 3): the internal force with
        (mu, lambda) = (1, 0) and (0, 1), since the stress is linear in both
        (ElasticEnergy::ComputeForceParamGrads, ElasticEnergy.cpp:284-315)"""
        x = self._x(x)
        out = []
        mu0, lam0 = self.mu.copy(), self.lam.copy()
        try:
            for a, b in ((1.0, 0.0), (0.0, 1.0)):
                self.mu[:] = a
                self.lam[:] = b
                out.append(self.internal_forces(x))
        finally:
            self.mu[:] = mu0
            self.lam[:] = lam0
        return out[0], out[1]

    def nodal_volumes(self):
        """sum over each node's elements of V / 4 (signed rest volumes)"""
        nv = np.zeros(self.nn)
        np.add.at(nv, self.tets.reshape(-1), np.repeat(0.25 * self.vol, 4))
        return nv

    def matvec(self, x, d):
        x = self._x(x)
        d = _f64(d).reshape(-1, 3)
        q = np.zeros_like(x)
        lib().orc_matvec(C.byref(self.s), _p(x), _p(d), _p(q))
        return q

    def pressure_forces(self, x=None, frac=1.0):
        x = self._x(x)
        f = np.zeros_like(x)
        lib().orc_pressure_forces(C.byref(self.s), _p(x), frac, _p(f))
        return f

    def external(self, x=None, frac=1.0):
        x = self._x(x)
        r = np.zeros_like(x)
        lib().orc_external(C.byref(self.s), _p(x), frac, _p(r))
        return r

    def gradients(self, x=None, frac=1.0):
        x = self._x(x)
        r = np.zeros_like(x)
        lib().orc_gradients(C.byref(self.s), _p(x), frac, _p(r))
        return r

    def block_diag(self, x=None):
        x = self._x(x)
        B = np.zeros((self.nn, 9))
        lib().orc_block_diag(C.byref(self.s), _p(x), _p(B))
        return B.reshape(-1, 3, 3)

    def cg(self, x, b, maxit, eps=1e-3, Binv=None, x0=None):
        x = self._x(x)
        b = _f64(b).reshape(-1, 3)
        sol = np.zeros_like(x) if x0 is None else _f64(x0).reshape(-1, 3).copy()
        Bi = None if Binv is None else _f64(Binv).reshape(-1, 9)
        k = lib().orc_cg(C.byref(self.s), _p(x), _p(b), _p(sol), int(maxit), float(eps), _p(Bi))
        return sol, k

    def newton_cg_ref(self, x=None, frac=1.0, outer=10, inner=100):
        x = self._x(x).copy()
        its = np.zeros(outer, dtype=np.int32)
        lib().orc_newton_cg_ref(C.byref(self.s), _p(x), frac, outer, inner, _p(its, _ip))
        return x, its

    def newton_solve(self, x=None, frac=1.0, tol=1e-10, max_newton=50, cg_rtol=1e-1, max_cg=100000):
        x = self._x(x).copy()
        st = np.zeros(4)
        rc = lib().orc_newton_solve(C.byref(self.s), _p(x), frac, tol, max_newton, cg_rtol, max_cg, _p(st))
        return x, {"converged": rc == 0, "newton": int(st[0]), "cg": int(st[1]),
                   "res": st[2], "ext": st[3]}


# ------------------------------------------------------- compiled reference
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.ref_create.argtypes = [_ip, C.c_int64, _dp, C.c_int64, C.POINTER(C.c_uint8), C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_int, C.c_int]
        L.ref_create.restype = C.c_void_p
        for name, args, res in [
            ("ref_destroy", [C.c_void_p], None),
            ("ref_set_positions", [C.c_void_p, _dp], None),
            ("ref_get_positions", [C.c_void_p, _dp], None),
            ("ref_set_pressure_surface", [C.c_void_p, _ip, C.c_int64, C.c_double], None),
            ("ref_gradients", [C.c_void_p, _dp], None),
            ("ref_matvec", [C.c_void_p, _dp, _dp], None),
            ("ref_cg", [C.c_void_p, _dp, _dp, C.c_int], C.c_int),
            ("ref_step", [C.c_void_p], None),
            ("ref_step_dt", [C.c_void_p, C.c_double], None),
            ("ref_newton_cg", [C.c_void_p, C.c_double], None),
            ("ref_set_force_fraction", [C.c_void_p, C.c_double], None),
            ("ref_force_fraction", [C.c_void_p], C.c_double),
            ("ref_energy", [C.c_void_p], C.c_double),
            ("ref_elastic_energy", [C.c_void_p], C.c_double),
            ("ref_spring_energy", [C.c_void_p], C.c_double),
            ("ref_add_cable", [C.c_void_p, _ip, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double], None),
            ("ref_set_cable_actuation", [C.c_void_p, C.c_int, C.c_double], None),
            ("ref_force_param_grads", [C.c_void_p, _dp, _dp, _dp], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _ref = L
    return _ref


class RefProblem:
    """The compiled reference FemPhysicsMatrixFree on the same inputs.
    Uniform material only (the reference has one mu, lambda); gravity along
    z when zup (FemPhysicsBase.cpp:30-33)."""

    ST_STATIC, ST_QUASI_STATIC, ST_EXPLICIT, ST_IMPLICIT = 0, 1, 2, 3

    def __init__(self, X0, tets, fixed, material, young, poisson, density=1070.0, gravity=-9.8,
                 zup=True, simtype=0, force_step=0.1, outer=10, inner=100, tris=None, pressure=0.0):
        self.X0 = _f64(X0).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        self.fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        L = reflib()
        # the reference prints its Lame parameters on construction (printf to
        # fd 1): keep stdout clean for callers that emit JSON
        import sys
        sys.stdout.flush()
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            self.h = L.ref_create(_p(self.tets, _ip), self.tets.shape[0], _p(self.X0), self.X0.shape[0],
                                  _p(self.fixed, C.POINTER(C.c_uint8)), int(material), float(young),
                                  float(poisson), float(density), float(gravity), int(zup), int(simtype),
                                  float(force_step), int(outer), int(inner))
            C.CDLL(None).fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.close(devnull)
        if tris is not None:
            self.tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
            L.ref_set_pressure_surface(self.h, _p(self.tris, _ip), self.tris.shape[0], float(pressure))

    def __del__(self):
        if getattr(self, "h", None):
            reflib().ref_destroy(self.h)
            self.h = None

    def set_positions(self, x):
        x = _f64(x).reshape(-1, 3)
        reflib().ref_set_positions(self.h, _p(x))

    def positions(self):
        x = np.zeros_like(self.X0)
        reflib().ref_get_positions(self.h, _p(x))
        return x

    def set_fraction(self, f):
        reflib().ref_set_force_fraction(self.h, float(f))

    def gradients(self):
        r = np.zeros_like(self.X0)
        reflib().ref_gradients(self.h, _p(r))
        return r

    def matvec(self, d):
        d = _f64(d).reshape(-1, 3)
        q = np.zeros_like(self.X0)
        reflib().ref_matvec(self.h, _p(d), _p(q))
        return q

    def cg(self, b, iters):
        b = _f64(b).reshape(-1, 3)
        x = np.zeros_like(self.X0)
        k = reflib().ref_cg(self.h, _p(b), _p(x), int(iters))
        return x, k

    def step(self):
        reflib().ref_step(self.h)

    def step_dt(self, dt):
        reflib().ref_step_dt(self.h, float(dt))

    def newton_cg(self, frac):
        reflib().ref_newton_cg(self.h, float(frac))

    def energy(self):
        return reflib().ref_energy(self.h)

    def elastic_energy(self):
        return reflib().ref_elastic_energy(self.h)

    def spring_energy(self):
        return reflib().ref_spring_energy(self.h)

    def add_cable(self, elem, bary, rest, stiffness=1000.0, damping=0.0, actuation=1.0):
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        bary = _f64(bary).reshape(-1, 3)
        reflib().ref_add_cable(self.h, _p(elem, _ip), _p(bary), len(elem), float(rest), float(stiffness),
                               float(damping), float(actuation))

    def set_cable_actuation(self, cable, actuation):
        reflib().ref_set_cable_actuation(self.h, int(cable), float(actuation))

    def force_param_grads(self):
        """GetForceParamGrads: (d f/d mu, d f/d lambda, d f/d rho) as (nn, 3), zero on fixed rows"""
        a, b, c = (np.zeros_like(self.X0) for _ in range(3))
        reflib().ref_force_param_grads(self.h, _p(a), _p(b), _p(c))
        return a, b, c

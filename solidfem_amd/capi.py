"""ctypes binding of libsfem_hip.so (the C-ABI in include/sfem_capi.h).

The shared library is the product path; there is no fallback.  Importing this
module does not need a GPU, but creating a context does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfem_hip.so")

LINEAR, STVK, COROTATIONAL, NEO_HOOKEAN = 0, 1, 2, 3
COROTATIONAL_POLAR = 16     # the polar rotation and its consistent tangent (no reference counterpart)
ST_STATIC, ST_QUASI_STATIC, ST_EXPLICIT, ST_IMPLICIT = 0, 1, 2, 3
SOLVE_REFERENCE, SOLVE_CONVERGED = 0, 1
PC_NONE, PC_JACOBI, PC_LEARNED, PC_MULTIGRID, PC_MULTIGRID_LEARNED = 0, 1, 2, 3, 4
MG_DENSE_MAX = 1536         # dofs of a coarsest multigrid level solved directly (SFEM_MG_DENSE_MAX)
ENOTCONV = -4

EXPORTS = [
    "sfem_create", "sfem_destroy", "sfem_last_error", "sfem_set_positions", "sfem_get_positions",
    "sfem_set_load_fraction", "sfem_get_load_fraction", "sfem_set_pressure_surface",
    "sfem_set_external_forces", "sfem_internal_forces", "sfem_element_stress", "sfem_get_velocities",
    "sfem_gradients", "sfem_matvec", "sfem_matvec_noalloc", "sfem_energy", "sfem_force_param_grads",
    "sfem_add_cable", "sfem_set_cable_actuation", "sfem_num_cables", "sfem_get_cable_positions", "sfem_clear_cables",
    "sfem_block_diagonal", "sfem_solve", "sfem_step", "sfem_set_precond_blocks", "sfem_info",
    "sfem_reset", "sfem_profile_hessvec", "sfem_time_hessvec",
    "sfem_mg_add_level", "sfem_mg_clear", "sfem_mg_info", "sfem_mg_apply", "sfem_mg_pcg", "sfem_mg_coarse_dense",
    "sfem_mg_set_schedule",
    "sfem_set_tangent_cache",
    "sfem_gnn_create", "sfem_gnn_destroy", "sfem_gnn_num_params", "sfem_gnn_get_params", "sfem_gnn_set_params",
    "sfem_set_learned_preconditioner", "sfem_collect_pairs", "sfem_collect_rough_pairs", "sfem_num_pairs", "sfem_get_pair", "sfem_clear_pairs",
    "sfem_gnn_loss", "sfem_gnn_eval", "sfem_gnn_train", "sfem_gnn_inputs", "sfem_gnn_gemm_check",
]

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Mesh(C.Structure):
    _fields_ = [("tets", _ip), ("ne", C.c_int64), ("X0", _dp), ("nn", C.c_int64),
                ("fixed", C.POINTER(C.c_uint8))]


class Material(C.Structure):
    _fields_ = [("material", C.c_int32), ("young", C.c_double), ("poisson", C.c_double),
                ("young_e", _dp), ("poisson_e", _dp), ("density", C.c_double),
                ("gravity", C.c_double * 3)]


class SolveParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("precond", C.c_int32), ("outer_iterations", C.c_int32),
                ("inner_iterations", C.c_int32), ("tol", C.c_double), ("cg_rtol", C.c_double),
                ("line_search", C.c_int32), ("check_every", C.c_int32), ("sim_type", C.c_int32),
                ("force_step", C.c_double), ("abs_tol", C.c_double), ("predictor", C.c_int32),
                ("time_step", C.c_double), ("substeps", C.c_int32), ("pc_newton_cap", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int32), ("cg_iters", C.c_int32), ("converged", C.c_int32),
                ("fallbacks", C.c_int32), ("res0", C.c_double), ("res", C.c_double),
                ("ext_norm", C.c_double), ("seconds", C.c_double), ("hessvec_evals", C.c_int64),
                ("force_evals", C.c_int64), ("reruns", C.c_int32), ("rejects", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sfem error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libsfem_hip.so; raises if it is not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        V = C.c_void_p
        sig = {
            "sfem_create": ([C.POINTER(Mesh), C.POINTER(Material), C.POINTER(V)], C.c_int),
            "sfem_destroy": ([V], None),
            "sfem_last_error": ([], C.c_char_p),
            "sfem_set_positions": ([V, _dp], C.c_int),
            "sfem_get_positions": ([V, _dp], C.c_int),
            "sfem_set_load_fraction": ([V, C.c_double], C.c_int),
            "sfem_get_load_fraction": ([V, _dp], C.c_int),
            "sfem_set_pressure_surface": ([V, _ip, C.c_int64, C.c_double], C.c_int),
            "sfem_set_external_forces": ([V, _dp], C.c_int),
            "sfem_internal_forces": ([V, _dp], C.c_int),
            "sfem_element_stress": ([V, _dp], C.c_int),
            "sfem_get_velocities": ([V, _dp], C.c_int),
            "sfem_gradients": ([V, _dp], C.c_int),
            "sfem_matvec": ([V, _dp, _dp], C.c_int),
            "sfem_matvec_noalloc": ([V, _dp, _dp], C.c_int),
            "sfem_energy": ([V, _dp], C.c_int),
            "sfem_force_param_grads": ([V, _dp, _dp, _dp], C.c_int),
            "sfem_add_cable": ([V, _ip, _dp, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32)],
                               C.c_int),
            "sfem_set_cable_actuation": ([V, C.c_int32, C.c_double], C.c_int),
            "sfem_num_cables": ([V, C.POINTER(C.c_int32)], C.c_int),
            "sfem_get_cable_positions": ([V, C.c_int32, _dp], C.c_int),
            "sfem_clear_cables": ([V], C.c_int),
            "sfem_block_diagonal": ([V, _dp], C.c_int),
            "sfem_solve": ([V, C.POINTER(SolveParams), C.POINTER(SolveStats)], C.c_int),
            "sfem_step": ([V, C.POINTER(SolveParams), C.POINTER(SolveStats)], C.c_int),
            "sfem_set_precond_blocks": ([V, _fp], C.c_int),
            "sfem_info": ([V, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                           C.POINTER(C.c_int64)], C.c_int),
            "sfem_reset": ([V], C.c_int),
            "sfem_profile_hessvec": ([V, C.c_int, _dp, _dp], C.c_int),
            "sfem_time_hessvec": ([V, C.c_int, _dp, C.POINTER(C.c_int64)], C.c_int),
            "sfem_mg_add_level": ([V, C.POINTER(Mesh), _ip, C.POINTER(C.c_int64), _ip, _dp, _dp, _dp], C.c_int),
            "sfem_mg_clear": ([V], C.c_int),
            "sfem_mg_info": ([V, C.POINTER(C.c_int32), _dp], C.c_int),
            "sfem_mg_coarse_dense": ([V, C.c_int64, C.POINTER(C.c_int32), _dp, _dp], C.c_int),
            "sfem_mg_apply": ([V, _dp, _dp], C.c_int),
            "sfem_mg_pcg": ([V, _dp, _dp, C.c_double, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
            "sfem_mg_set_schedule": ([V, C.c_int], C.c_int),
            "sfem_set_tangent_cache": ([V, C.c_int], C.c_int),
            "sfem_gnn_create": ([C.c_uint64, C.POINTER(V)], C.c_int),
            "sfem_gnn_destroy": ([V], None),
            "sfem_gnn_num_params": ([V, C.POINTER(C.c_int64)], C.c_int),
            "sfem_gnn_get_params": ([V, _fp], C.c_int),
            "sfem_gnn_set_params": ([V, _fp], C.c_int),
            "sfem_set_learned_preconditioner": ([V, V], C.c_int),
            "sfem_collect_pairs": ([V, _dp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int)], C.c_int),
            "sfem_collect_rough_pairs": ([V, C.c_int, C.c_uint64, C.POINTER(C.c_int)], C.c_int),
            "sfem_num_pairs": ([V, C.POINTER(C.c_int64)], C.c_int),
            "sfem_get_pair": ([V, C.c_int64, _fp, _dp], C.c_int),
            "sfem_clear_pairs": ([V], C.c_int),
            "sfem_gnn_loss": ([V, V, C.c_int64, C.c_int, _dp, _fp], C.c_int),
            "sfem_gnn_eval": ([V, V, _dp, _dp], C.c_int),
            "sfem_gnn_train": ([V, V, C.c_int, C.c_int, C.c_float, C.c_uint64, _dp], C.c_int),
            "sfem_gnn_inputs": ([V, V, _fp, _fp, _fp], C.c_int),
            "sfem_gnn_gemm_check": ([V, C.c_int, C.c_int64, _fp, _fp], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc, allow=()):
    if rc != 0 and rc not in allow:
        raise SfemError(rc, lib().sfem_last_error().decode())
    return rc


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def solve_params(mode=SOLVE_CONVERGED, precond=PC_JACOBI, outer=50, inner=20000, tol=1e-8, cg_rtol=0.0,
                 line_search=1, check_every=16, sim_type=ST_STATIC, force_step=0.1, abs_tol=0.0, predictor=2,
                 time_step=0.016, substeps=1, pc_newton_cap=0):
    return SolveParams(mode, precond, outer, inner, tol, cg_rtol, line_search, check_every, sim_type, force_step,
                       abs_tol, predictor, time_step, substeps, pc_newton_cap)


class Solver:
    """One device context: mesh, material, loads, state (original node order)."""

    def __init__(self, X0, tets, fixed, material, young=1e6, poisson=0.45, density=1070.0,
                 gravity=(0.0, 0.0, 0.0), young_e=None, poisson_e=None):
        self.X0 = np.ascontiguousarray(X0, dtype=np.float64).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        self.fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        self.nn, self.ne = self.X0.shape[0], self.tets.shape[0]
        self._ye = None if young_e is None else np.ascontiguousarray(young_e, dtype=np.float64)
        self._pe = None if poisson_e is None else np.ascontiguousarray(poisson_e, dtype=np.float64)
        mesh = Mesh(_p(self.tets, _ip), self.ne, _p(self.X0), self.nn, _p(self.fixed, C.POINTER(C.c_uint8)))
        g = (C.c_double * 3)(*gravity)
        mat = Material(int(material), float(young), float(poisson), _p(self._ye), _p(self._pe), float(density), g)
        h = C.c_void_p()
        _check(lib().sfem_create(C.byref(mesh), C.byref(mat), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sfem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        v = [C.c_int64() for _ in range(4)]
        _check(lib().sfem_info(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("ne", "nn", "nfixed", "nblocks"), (x.value for x in v)))

    def reset(self):
        _check(lib().sfem_reset(self.h))

    def profile_hessvec(self, reps=100):
        a, b = C.c_double(), C.c_double()
        _check(lib().sfem_profile_hessvec(self.h, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_hessvec(self, every):
        """start (every > 0) / stop (0) sampled event timing of the CG loops'
        Hessian-vector launches; -> (average ms, samples) since the last start"""
        a, n = C.c_double(), C.c_int64()
        _check(lib().sfem_time_hessvec(self.h, int(every), C.byref(a), C.byref(n)))
        return a.value, n.value

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        _check(lib().sfem_set_positions(self.h, _p(x)))

    def positions(self):
        x = np.zeros((self.nn, 3))
        _check(lib().sfem_get_positions(self.h, _p(x)))
        return x

    def set_load_fraction(self, f):
        _check(lib().sfem_set_load_fraction(self.h, float(f)))

    def load_fraction(self):
        f = C.c_double()
        _check(lib().sfem_get_load_fraction(self.h, C.byref(f)))
        return f.value

    def set_pressure_surface(self, tris, pressure):
        self._tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        _check(lib().sfem_set_pressure_surface(self.h, _p(self._tris, _ip), self._tris.shape[0], float(pressure)))

    def set_external_forces(self, f):
        self._fext = None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        _check(lib().sfem_set_external_forces(self.h, _p(self._fext)))

    def internal_forces(self):
        f = np.zeros((self.nn, 3))
        _check(lib().sfem_internal_forces(self.h, _p(f)))
        return f

    def element_stress(self):
        """first Piola-Kirchhoff stress per element (ne, 3, 3), original element order"""
        P = np.zeros((self.ne, 3, 3))
        _check(lib().sfem_element_stress(self.h, _p(P)))
        return P

    def velocities(self):
        v = np.zeros((self.nn, 3))
        _check(lib().sfem_get_velocities(self.h, _p(v)))
        return v

    def gradients(self):
        r = np.zeros((self.nn, 3))
        _check(lib().sfem_gradients(self.h, _p(r)))
        return r

    def matvec(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        q = np.zeros_like(d)
        _check(lib().sfem_matvec(self.h, _p(d), _p(q)))
        return q

    def matvec_noalloc(self, d):
        """sfem_matvec_noalloc: K d with no allocation (static per-device scratch)"""
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        q = np.zeros_like(d)
        _check(lib().sfem_matvec_noalloc(self.h, _p(d), _p(q)))
        return q

    def energy(self):
        """(elastic energy, cable spring energy, work of the dead loads) at the current state (sfem_energy)"""
        out = np.zeros(3)
        _check(lib().sfem_energy(self.h, _p(out)))
        return tuple(out)

    def force_param_grads(self):
        """(d f/d mu, d f/d lambda, d (m g)/d rho), each (nn, 3), original order"""
        a, b, c = (np.zeros((self.nn, 3)) for _ in range(3))
        _check(lib().sfem_force_param_grads(self.h, _p(a), _p(b), _p(c)))
        return a, b, c

    def add_cable(self, elem, bary, rest_length, stiffness=1000.0, damping=0.0):
        """a cable of len(elem) nodes embedded in elements elem at barycentric weights bary (n, 3) -> its index"""
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        bary = np.ascontiguousarray(bary, dtype=np.float64).reshape(-1, 3)
        i = C.c_int32()
        _check(lib().sfem_add_cable(self.h, _p(elem, _ip), _p(bary), len(elem), float(rest_length), float(stiffness),
                                    float(damping), C.byref(i)))
        return i.value

    def set_cable_actuation(self, cable, actuation):
        _check(lib().sfem_set_cable_actuation(self.h, int(cable), float(actuation)))

    def num_cables(self):
        n = C.c_int32()
        _check(lib().sfem_num_cables(self.h, C.byref(n)))
        return n.value

    def cable_positions(self, cable, n):
        pos = np.zeros((n, 3))
        _check(lib().sfem_get_cable_positions(self.h, int(cable), _p(pos)))
        return pos

    def clear_cables(self):
        _check(lib().sfem_clear_cables(self.h))

    def block_diagonal(self):
        B = np.zeros((self.nn, 9))
        _check(lib().sfem_block_diagonal(self.h, _p(B)))
        return B.reshape(-1, 3, 3)

    def set_precond_blocks(self, minv):
        self._minv = None if minv is None else np.ascontiguousarray(minv, dtype=np.float32).reshape(-1, 9)
        _check(lib().sfem_set_precond_blocks(self.h, _p(self._minv, _fp)))

    def add_mg_level(self, X0, tets, fixed, inject, p_ptr, p_col, p_val, young_e=None, poisson_e=None):
        """one coarser level of the multigrid preconditioner (sfem_mg_add_level)"""
        ye = None if young_e is None else np.ascontiguousarray(young_e, dtype=np.float64)
        pe = None if poisson_e is None else np.ascontiguousarray(poisson_e, dtype=np.float64)
        X0 = np.ascontiguousarray(X0, dtype=np.float64).reshape(-1, 3)
        tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        inject = np.ascontiguousarray(inject, dtype=np.int32)
        p_ptr = np.ascontiguousarray(p_ptr, dtype=np.int64)
        p_col = np.ascontiguousarray(p_col, dtype=np.int32)
        p_val = np.ascontiguousarray(p_val, dtype=np.float64)
        m = Mesh(_p(tets, _ip), tets.shape[0], _p(X0), X0.shape[0], fixed.ctypes.data_as(C.POINTER(C.c_uint8)))
        _check(lib().sfem_mg_add_level(self.h, C.byref(m), _p(inject, _ip), p_ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                       _p(p_col, _ip), _p(p_val), _p(ye), _p(pe)))

    def set_multigrid(self, levels):
        """levels: [(X0, tets, fixed, inject, p_ptr, p_col, p_val[, young_e, poisson_e])], finest
        coarse level first (solidfem_amd.mesh.beam_hierarchy)"""
        _check(lib().sfem_mg_clear(self.h))
        for lv in levels:
            self.add_mg_level(*lv)

    def mg_info(self):
        n = C.c_int32()
        _check(lib().sfem_mg_info(self.h, C.byref(n), None))
        lm = np.zeros(n.value)
        _check(lib().sfem_mg_info(self.h, C.byref(n), _p(lm)))
        return lm

    def mg_coarse_dense(self):
        """(A, Ainv) of the directly solved coarsest level at the current state"""
        n = C.c_int32()
        _check(lib().sfem_mg_coarse_dense(self.h, 0, C.byref(n), None, None))
        a = np.zeros((n.value, n.value))
        ai = np.zeros((n.value, n.value))
        _check(lib().sfem_mg_coarse_dense(self.h, a.size, C.byref(n), _p(a), _p(ai)))
        return a, ai

    def mg_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        z = np.zeros_like(r)
        _check(lib().sfem_mg_apply(self.h, _p(r), _p(z)))
        return z

    def mg_set_schedule(self, fused):
        """the V-cycle's launch schedule: True folds the smoothing steps into the
        element kernels (the default), False launches one kernel per step"""
        _check(lib().sfem_mg_set_schedule(self.h, 1 if fused else 0))

    def set_tangent_cache(self, mode):
        """Neo-Hookean static solves: 1 (the default) builds the tangent's
        state per element once per Newton step and makes that step's products
        from it; 0 rebuilds it in every product; 2 as 1, and matvec() builds
        it at each call too"""
        _check(lib().sfem_set_tangent_cache(self.h, int(mode)))

    def mg_pcg(self, b, rtol=1e-6, maxit=500):
        """-> (x, iterations, status) of one multigrid-PCG solve K x = b"""
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
        x = np.zeros_like(b)
        it, stt = C.c_int32(), C.c_int32()
        _check(lib().sfem_mg_pcg(self.h, _p(b), _p(x), rtol, maxit, C.byref(it), C.byref(stt)))
        return x, it.value, stt.value

    def solve(self, params=None, allow_notconv=False, **kw):
        p = params if params is not None else solve_params(**kw)
        st = SolveStats()
        _check(lib().sfem_solve(self.h, C.byref(p), C.byref(st)), allow=(ENOTCONV,) if allow_notconv else ())
        return st.as_dict()

    def step(self, params=None, allow_notconv=False, **kw):
        p = params if params is not None else solve_params(**kw)
        st = SolveStats()
        _check(lib().sfem_step(self.h, C.byref(p), C.byref(st)), allow=(ENOTCONV,) if allow_notconv else ())
        return st.as_dict()


class LearnedPreconditioner:
    """The message-passing network that predicts per-node 3x3 SPD blocks
    (sfem_gnn_* in include/sfem_capi.h); weights live on the device."""

    LAYERS = [(16, 32), (32, 32), (64, 32), (32, 32), (64, 32), (32, 32), (64, 32), (32, 32), (32, 32), (32, 16)]

    def __init__(self, seed=1):
        h = C.c_void_p()
        _check(lib().sfem_gnn_create(int(seed), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sfem_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_params(self):
        n = C.c_int64()
        _check(lib().sfem_gnn_num_params(self.h, C.byref(n)))
        return n.value

    def params(self):
        w = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().sfem_gnn_get_params(self.h, _p(w, _fp)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.num_params()
        _check(lib().sfem_gnn_set_params(self.h, _p(w, _fp)))

    def attach(self, solver):
        """the context keeps a pointer to the network: the solver object holds
        a reference to it until detach, so the network outlives its use"""
        _check(lib().sfem_set_learned_preconditioner(solver.h, self.h))
        solver._learned = self

    @staticmethod
    def detach(solver):
        _check(lib().sfem_set_learned_preconditioner(solver.h, None))
        solver._learned = None

    @staticmethod
    def collect(solver, load, nsnap=8, snap_every=10, rtol=1e-6):
        load = np.ascontiguousarray(load, dtype=np.float64).reshape(-1, 3)
        n = C.c_int()
        _check(lib().sfem_collect_pairs(solver.h, _p(load), int(nsnap), int(snap_every), float(rtol), C.byref(n)))
        return n.value

    @staticmethod
    def collect_rough(solver, count, seed=0):
        """count (K e, e.Ke) pairs with white-noise e at the solver's state"""
        n = C.c_int()
        _check(lib().sfem_collect_rough_pairs(solver.h, int(count), int(seed), C.byref(n)))
        return n.value

    @staticmethod
    def num_pairs(solver):
        n = C.c_int64()
        _check(lib().sfem_num_pairs(solver.h, C.byref(n)))
        return n.value

    @staticmethod
    def pair(solver, k):
        r = np.zeros((solver.nn, 3), dtype=np.float32)
        er = C.c_double()
        _check(lib().sfem_get_pair(solver.h, int(k), _p(r, _fp), C.byref(er)))
        return r, er.value

    @staticmethod
    def clear_pairs(solver):
        _check(lib().sfem_clear_pairs(solver.h))

    def loss(self, solver, first, count, grad=False):
        l = C.c_double()
        g = np.zeros(self.num_params(), dtype=np.float32) if grad else None
        _check(lib().sfem_gnn_loss(self.h, solver.h, int(first), int(count), C.byref(l), _p(g, _fp)))
        return (l.value, g) if grad else l.value

    def evaluate(self, solver):
        a, b = C.c_double(), C.c_double()
        _check(lib().sfem_gnn_eval(self.h, solver.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def train(self, solver, steps, batch=8, lr=1e-3, seed=0):
        losses = np.zeros(steps)
        _check(lib().sfem_gnn_train(self.h, solver.h, int(steps), int(batch), float(lr), int(seed), _p(losses)))
        return losses

    def inputs(self, solver):
        f = np.zeros((solver.nn, 16), dtype=np.float32)
        c = np.zeros((solver.nn, 6), dtype=np.float32)
        m = np.zeros((solver.nn, 9), dtype=np.float32)
        _check(lib().sfem_gnn_inputs(self.h, solver.h, _p(f, _fp), _p(c, _fp), _p(m, _fp)))
        return f, c, m

    def gemm_check(self, layer, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        K, NO = self.LAYERS[layer]
        assert X.shape[1] == K
        Y = np.zeros((X.shape[0], NO), dtype=np.float32)
        _check(lib().sfem_gnn_gemm_check(self.h, int(layer), X.shape[0], _p(X, _fp), _p(Y, _fp)))
        return Y

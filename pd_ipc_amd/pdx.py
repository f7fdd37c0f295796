"""ctypes host mirror of the C ABI in include/pdx.h.

The names follow the reference's interface: ``Engine.step_once`` is
``pd::stepOnce`` (core/sim/Integrator.h:14), ``Engine.solve`` is
``EigenDirectSolver::solve`` (core/solver/EigenDirectSolver.cpp:299-357), and
``Engine.stats`` carries the reference's counters (core/sim/Scene.h:161-169,
core/solver/IGlobalSolver.h:114-125).  Errors raise ``PdxError`` where the
reference aborts.  The library must be built (``python -m pd_ipc_amd.build``);
there is no fallback.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdx.so")

PDX_PREC_FP64 = 0
PDX_PREC_MIXED32 = 1
PDX_SOLVE_EXACT = 0
PDX_SOLVE_FAST = 1
PDX_SOLVE_DENSE = 2
PDX_ORDER_AMD = 0
PDX_ORDER_ND = 1

_STATUS = {0: "ok", -1: "invalid argument", -2: "not ready", -3: "not positive definite",
           -4: "HIP runtime error", -5: "unsupported", -6: "out of memory"}


class PdxError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"pdx error {code} ({_STATUS.get(code, '?')}): {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("damping", C.c_double),
                ("max_iterations", C.c_int), ("rel_tol", C.c_double), ("abs_tol", C.c_double),
                ("residual_tol", C.c_double), ("chebyshev_rho", C.c_double),
                ("chebyshev_start", C.c_int)]


class _Scene(C.Structure):
    _fields_ = [("n_vertices", C.c_int), ("positions", C.c_void_p), ("velocities", C.c_void_p),
                ("masses", C.c_void_p), ("pinned", C.c_void_p), ("pin_positions", C.c_void_p),
                ("n_edges", C.c_int), ("edge_a", C.c_void_p), ("edge_b", C.c_void_p),
                ("edge_rest", C.c_void_p), ("edge_kappa", C.c_void_p), ("n_bends", C.c_int),
                ("bend_a", C.c_void_p), ("bend_b", C.c_void_p), ("bend_c", C.c_void_p),
                ("bend_k", C.c_void_p)]


class _Stats(C.Structure):
    _fields_ = [("iterations_used", C.c_int), ("converged", C.c_int), ("early_exit_count", C.c_int),
                ("factorize_count", C.c_int), ("analyze_count", C.c_int), ("color_count", C.c_int),
                ("total_iterations", C.c_longlong), ("substeps", C.c_longlong),
                ("last_residual", C.c_double), ("factor_nnz", C.c_longlong),
                ("fwd_levels", C.c_int), ("bwd_levels", C.c_int), ("fwd_steps", C.c_int),
                ("bwd_steps", C.c_int), ("analyze_seconds", C.c_double),
                ("factorize_seconds", C.c_double), ("phase_seconds", C.c_double * 9),
                ("workgroup_seconds", C.c_double * 4)]


class _Diag(C.Structure):
    _fields_ = [("max_strain", C.c_double), ("mean_strain", C.c_double), ("elastic_energy", C.c_double),
                ("total_energy", C.c_double), ("max_speed", C.c_double), ("total_mass", C.c_double)]


class _BenchSummary(C.Structure):
    _fields_ = [("n_vertices", C.c_int), ("n_edges", C.c_int), ("n_bends", C.c_int), ("n_pinned", C.c_int),
                ("substeps", C.c_int), ("converged_substeps", C.c_int), ("max_iterations_used", C.c_int),
                ("total_iterations", C.c_longlong), ("last_residual", C.c_double),
                ("max_strain", C.c_double), ("mean_strain", C.c_double), ("run_max_strain", C.c_double),
                ("elastic_energy", C.c_double), ("total_energy", C.c_double), ("max_speed", C.c_double),
                ("total_mass", C.c_double), ("wall_seconds", C.c_double),
                ("device_ms_per_substep", C.c_double), ("analyze_count", C.c_int),
                ("factorize_count", C.c_int), ("color_count", C.c_int), ("matrix_nnz", C.c_longlong),
                ("factor_nnz", C.c_longlong), ("analyze_seconds", C.c_double),
                ("factorize_seconds", C.c_double), ("loaded_mesh", C.c_int)]


_lib = None


def load_library(path=LIB_PATH):
    """Loads libpdx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PdxError(-2, f"{path} is missing: run `python -m pd_ipc_amd.build`")
    lib = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    sig = {
        "pdx_create": (I, [C.POINTER(P), I, I, I]),
        "pdx_destroy": (I, [P]),
        "pdx_last_error": (C.c_char_p, [P]),
        "pdx_add_instance": (I, [P, C.POINTER(_Scene), C.POINTER(_Config), C.POINTER(I)]),
        "pdx_instance_count": (I, [P]),
        "pdx_factor": (I, [P]),
        "pdx_step": (I, [P, I]),
        "pdx_step_async": (I, [P, I]),
        "pdx_sync": (I, [P]),
        "pdx_step_graph": (I, [P, I, I]),
        "pdx_get_state": (I, [P, I, P, P]),
        "pdx_set_state": (I, [P, I, P, P]),
        "pdx_set_pin_positions": (I, [P, I, P]),
        "pdx_set_edge_kappa": (I, [P, I, P]),
        "pdx_set_bend_k": (I, [P, I, P]),
        "pdx_set_masses": (I, [P, I, P]),
        "pdx_set_pinned": (I, [P, I, P]),
        "pdx_set_config": (I, [P, I, C.POINTER(_Config)]),
        "pdx_get_stats": (I, [P, I, C.POINTER(_Stats)]),
        "pdx_solve3": (I, [P, I, P, P]),
        "pdx_solve3_noalloc": (I, [P, I, P, P]),
        "pdx_get_factor": (I, [P, I, C.POINTER(I), C.POINTER(C.c_longlong), P, P, P, P, P, P]),
        "pdx_get_topology": (I, [P, I, C.POINTER(I), P, P, P, P, P, P, P]),
        "pdx_pin_count": (I, [P, I]),
        "pdx_set_pin_trajectory": (I, [P, I, I, P]),
        "pdx_rollout": (I, [P, I, I]),
        "pdx_set_adjoint_seed": (I, [P, I, P, P]),
        "pdx_adjoint": (I, [P]),
        "pdx_get_gradients": (I, [P, I, P, P, P, P, P]),
        "pdx_rollout_timing": (I, [P, C.POINTER(D), C.POINTER(D)]),
        "pdx_last_kernel_ms": (D, [P]),
        "pdx_device_info": (I, [I, P, P, P, P]),
        "pdx_autotune": (I, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(D)]),
        "pdx_set_launch": (I, [P, I, I]),
        "pdx_get_schedule_stats": (I, [P, I, P]),
        "pdx_set_schedule": (I, [P, I, I]),
        "pdx_set_ordering": (I, [P, I]),
        "pdx_set_split": (I, [P, C.c_longlong]),
        "pdx_split_count": (I, [P]),
        "pdx_get_split": (C.c_longlong, [P]),
        "pdx_get_launch": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "pdx_verify_schedule": (I, [P, I, C.POINTER(D)]),
        "pdx_bytes_per_iteration": (D, [P]),
        "pdx_bytes_per_substep": (D, [P]),
        "pdx_grid_counts": (I, [I, I, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "pdx_grid_fill": (I, [I, I, D, D, D, D, D, P, P, P, P, P, P, P, P, P]),
        "pdx_obj_counts": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "pdx_obj_fill": (I, [C.c_char_p, D, P, P, P, P, P, P]),
        "pdx_masses_from_surfaces": (I, [I, P, I, P, D, P]),
        "pdx_masses_from_edges": (I, [I, I, P, P, P, D, P]),
        "pdx_diagnostics": (I, [P, I, C.POINTER(_Diag)]),
        "pdx_matrix_nnz": (I, [P, I, C.POINTER(C.c_longlong)]),
        "pdx_bench_run": (I, [I, C.POINTER(C.c_char_p), C.c_char_p, C.c_longlong, C.POINTER(_BenchSummary)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _arr(a, dtype, size, what):
    """`a` as a contiguous 1-D array of `size` values, or ValueError: the C side
    copies exactly that many values, so a short array would be read past its end."""
    out = np.ascontiguousarray(a, dtype).reshape(-1)
    if out.size != size:
        raise ValueError(f"{what}: expected {size} values, got {out.size}")
    return out


def device_info(device=0):
    """The device's compute units, shader clock (kHz), memory clock (kHz) and
    memory bus width (bits) as the HIP runtime states them."""
    lib = load_library()
    v = [C.c_int() for _ in range(4)]
    rc = lib.pdx_device_info(device, *[C.byref(x) for x in v])
    if rc != 0:
        raise PdxError(rc, "device_info")
    return dict(compute_units=v[0].value, clock_khz=v[1].value, mem_clock_khz=v[2].value,
                bus_width_bits=v[3].value)


@dataclass
class SimConfig:
    """The SceneConfig fields the substep reads (core/sim/Scene.h:14-137)."""
    dt: float = 1.0 / 120.0
    gravity: tuple = (0.0, -9.81, 0.0)
    damping: float = 0.02
    max_iterations: int = 10
    rel_tol: float = 1e-3
    abs_tol: float = 1e-12
    residual_tol: float = 1e-3
    chebyshev_rho: float = 0.0   # > 0 turns the Chebyshev blend on (reference PD_CHEB_RHO, 0.98)
    chebyshev_start: int = 10    # first blended iteration is start + 1 (reference PD_CHEB_START)

    @staticmethod
    def fixed_iterations(iters, dt, damping=0.02, gravity=(0.0, -9.81, 0.0)):
        """Fixed PD iteration count with the residual still evaluated every
        iteration (the reference's own setting, _verify/iteration_spectrum.cpp:106-108)."""
        return SimConfig(dt=dt, gravity=gravity, damping=damping, max_iterations=iters,
                         rel_tol=0.0, abs_tol=0.0, residual_tol=1e-30)

    def _c(self):
        c = _Config()
        c.dt = self.dt
        for i in range(3):
            c.gravity[i] = self.gravity[i]
        c.damping = self.damping
        c.max_iterations = self.max_iterations
        c.rel_tol = self.rel_tol
        c.abs_tol = self.abs_tol
        c.residual_tol = self.residual_tol
        c.chebyshev_rho = self.chebyshev_rho
        c.chebyshev_start = self.chebyshev_start
        return c


@dataclass
class Cloth:
    """One cloth instance as flat numpy arrays (the Mesh fields of core/mesh/Mesh.h)."""
    positions: np.ndarray
    masses: np.ndarray
    pinned: np.ndarray
    edge_a: np.ndarray
    edge_b: np.ndarray
    edge_rest: np.ndarray
    edge_kappa: np.ndarray
    velocities: np.ndarray = None
    pin_positions: np.ndarray = None
    bend_a: np.ndarray = None
    bend_b: np.ndarray = None
    bend_c: np.ndarray = None
    bend_k: np.ndarray = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        f8 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        i4 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        self.positions = f8(self.positions)
        self.masses = f8(self.masses)
        self.pinned = np.ascontiguousarray(self.pinned, dtype=np.uint8).reshape(-1)
        self.edge_a, self.edge_b = i4(self.edge_a), i4(self.edge_b)
        self.edge_rest, self.edge_kappa = f8(self.edge_rest), f8(self.edge_kappa)
        self.velocities = f8(self.velocities) if self.velocities is not None else np.zeros_like(self.positions)
        self.pin_positions = f8(self.pin_positions) if self.pin_positions is not None else self.positions.copy()
        if self.bend_a is None or len(self.bend_a) == 0:
            self.bend_a = self.bend_b = self.bend_c = np.zeros(0, np.int32)
            self.bend_k = np.zeros(0, np.float64)
        else:
            self.bend_a, self.bend_b, self.bend_c = i4(self.bend_a), i4(self.bend_b), i4(self.bend_c)
            self.bend_k = f8(self.bend_k)

    @property
    def n(self):
        return len(self.masses)

    @property
    def n_edges(self):
        return len(self.edge_a)

    @property
    def n_bends(self):
        return len(self.bend_a)

    def _c(self):
        s = _Scene()
        s.n_vertices = self.n
        s.positions, s.velocities = _ptr(self.positions), _ptr(self.velocities)
        s.masses, s.pinned, s.pin_positions = _ptr(self.masses), _ptr(self.pinned), _ptr(self.pin_positions)
        s.n_edges = self.n_edges
        s.edge_a, s.edge_b = _ptr(self.edge_a), _ptr(self.edge_b)
        s.edge_rest, s.edge_kappa = _ptr(self.edge_rest), _ptr(self.edge_kappa)
        s.n_bends = self.n_bends
        s.bend_a, s.bend_b, s.bend_c, s.bend_k = (_ptr(self.bend_a), _ptr(self.bend_b),
                                                  _ptr(self.bend_c), _ptr(self.bend_k))
        return s


class Engine:
    """A context of one or more cloth instances stepped together on one GPU."""

    def __init__(self, precision="fp64", solve="exact", device=0):
        self.lib = load_library()
        prec = {"fp64": PDX_PREC_FP64, "mixed32": PDX_PREC_MIXED32}[precision]
        mode = {"exact": PDX_SOLVE_EXACT, "fast": PDX_SOLVE_FAST, "dense": PDX_SOLVE_DENSE}[solve]
        h = C.c_void_p()
        rc = self.lib.pdx_create(C.byref(h), device, prec, mode)
        if rc != 0:
            raise PdxError(rc, "pdx_create")
        self.h = h
        self.precision = precision
        self.cloths = []

    def close(self):
        if self.h:
            self.lib.pdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise PdxError(rc, f"{what}: {self.lib.pdx_last_error(self.h).decode(errors='replace')}")

    def add(self, cloth, cfg):
        idx = C.c_int()
        s, c = cloth._c(), cfg._c()
        self._check(self.lib.pdx_add_instance(self.h, C.byref(s), C.byref(c), C.byref(idx)), "add_instance")
        self.cloths.append(cloth)
        return idx.value

    def factor(self):
        self._check(self.lib.pdx_factor(self.h), "factor")

    def step(self, nsub=1):
        self._check(self.lib.pdx_step(self.h, nsub), "step")

    def step_once(self):
        """pd::stepOnce on every instance of the context."""
        self.step(1)

    def step_async(self, nsub):
        self._check(self.lib.pdx_step_async(self.h, nsub), "step_async")

    def sync(self):
        self._check(self.lib.pdx_sync(self.h), "sync")

    def step_graph(self, nsub, replays):
        self._check(self.lib.pdx_step_graph(self.h, nsub, replays), "step_graph")

    def state(self, i=0):
        n = self.cloths[i].n
        x = np.empty(3 * n)
        v = np.empty(3 * n)
        self._check(self.lib.pdx_get_state(self.h, i, _ptr(x), _ptr(v)), "get_state")
        return x, v

    def _cloth(self, i):
        if not 0 <= i < len(self.cloths):
            raise IndexError(f"instance {i} of {len(self.cloths)}")
        return self.cloths[i]

    def set_state(self, i, x=None, v=None):
        n3 = 3 * self._cloth(i).n
        x = None if x is None else _arr(x, np.float64, n3, "set_state positions")
        v = None if v is None else _arr(v, np.float64, n3, "set_state velocities")
        self._check(self.lib.pdx_set_state(self.h, i, _ptr(x), _ptr(v)), "set_state")

    def set_pin_positions(self, i, q):
        q = _arr(q, np.float64, 3 * self._cloth(i).n, "set_pin_positions")
        self._check(self.lib.pdx_set_pin_positions(self.h, i, _ptr(q)), "set_pin_positions")

    def set_edge_kappa(self, i, kappa):
        k = _arr(kappa, np.float64, self._cloth(i).n_edges, "set_edge_kappa")
        self._check(self.lib.pdx_set_edge_kappa(self.h, i, _ptr(k)), "set_edge_kappa")

    def set_bend_k(self, i, k):
        k = _arr(k, np.float64, self._cloth(i).n_bends, "set_bend_k")
        self._check(self.lib.pdx_set_bend_k(self.h, i, _ptr(k)), "set_bend_k")

    def set_masses(self, i, masses):
        """Mesh::masses edit: the next step refactors once (no new analysis)."""
        m = _arr(masses, np.float64, self._cloth(i).n, "set_masses")
        self._check(self.lib.pdx_set_masses(self.h, i, _ptr(m)), "set_masses")

    def set_pinned(self, i, pinned):
        """Mesh::pinned edit: the next step analyses and factors again."""
        p = _arr(np.asarray(pinned) != 0, np.uint8, self._cloth(i).n, "set_pinned")
        self._check(self.lib.pdx_set_pinned(self.h, i, _ptr(p)), "set_pinned")

    def set_config(self, i, cfg):
        c = cfg._c()
        self._check(self.lib.pdx_set_config(self.h, i, C.byref(c)), "set_config")

    def stats(self, i=0):
        s = _Stats()
        self._check(self.lib.pdx_get_stats(self.h, i, C.byref(s)), "get_stats")
        out = {k: getattr(s, k) for k, _ in _Stats._fields_}
        out["phase_seconds"] = list(s.phase_seconds)
        out["workgroup_seconds"] = list(s.workgroup_seconds)
        return out

    def diagnostics(self, i=0):
        """Strain, energy and speed of the instance's current state, in the reference's
        summation order: Mesh::maxRelativeStrain / meanRelativeStrain / totalMass
        (core/mesh/Mesh.cpp:345-372), surrogateEnergy / totalEnergy / maxSpeed
        (core/sim/Integrator.cpp:804-846)."""
        d = _Diag()
        self._check(self.lib.pdx_diagnostics(self.h, i, C.byref(d)), "diagnostics")
        return {k: getattr(d, k) for k, _ in _Diag._fields_}

    def matrix_nnz(self, i=0):
        """Structural nonzeros of the 3n x 3n matrix L (IGlobalSolver::Stats::nnz)."""
        v = C.c_longlong()
        self._check(self.lib.pdx_matrix_nnz(self.h, i, C.byref(v)), "matrix_nnz")
        return v.value

    def solve(self, i, b):
        b = _arr(b, np.float64, 3 * self._cloth(i).n, "solve")
        x = np.empty_like(b)
        self._check(self.lib.pdx_solve3(self.h, i, _ptr(b), _ptr(x)), "solve3")
        return x

    def factor_arrays(self, i=0):
        m, nnz = C.c_int(), C.c_longlong()
        self._check(self.lib.pdx_get_factor(self.h, i, C.byref(m), C.byref(nnz), None, None, None,
                                            None, None, None), "get_factor")
        m, nnz = m.value, nnz.value
        out = dict(ptr=np.empty(m + 1, np.int32), rows=np.empty(nnz, np.int32), vals=np.empty(nnz),
                   dinv=np.empty(m), rank_vertex=np.empty(m, np.int32), perm3=np.empty(3 * m, np.int32))
        self._check(self.lib.pdx_get_factor(self.h, i, C.byref(C.c_int()), C.byref(C.c_longlong()),
                                            _ptr(out["ptr"]), _ptr(out["rows"]), _ptr(out["vals"]),
                                            _ptr(out["dinv"]), _ptr(out["rank_vertex"]),
                                            _ptr(out["perm3"])), "get_factor")
        return out

    def topology(self, i=0):
        """Edge colouring and incidence lists of instance i (Mesh::constraintColoring,
        Mesh::bendAdjacency)."""
        c = self.cloths[i]
        ncol = C.c_int()
        self._check(self.lib.pdx_get_topology(self.h, i, C.byref(ncol), None, None, None, None, None,
                                              None, None), "get_topology")
        n, E, B = c.n, len(c.edge_a), len(c.bend_a)
        out = dict(color_of=np.empty(E, np.uint32), order=np.empty(E, np.uint32),
                   start=np.empty(ncol.value + 1, np.uint32), vertex_start=np.empty(n + 1, np.uint32),
                   vertex_edges=np.empty(2 * E, np.uint32), bend_vertex_start=np.empty(n + 1, np.uint32),
                   bend_stencils=np.empty(3 * B, np.uint32))
        self._check(self.lib.pdx_get_topology(
            self.h, i, C.byref(ncol), _ptr(out["color_of"]), _ptr(out["order"]), _ptr(out["start"]),
            _ptr(out["vertex_start"]), _ptr(out["vertex_edges"]),
            _ptr(out["bend_vertex_start"]) if B else None, _ptr(out["bend_stencils"]) if B else None),
            "get_topology")
        out["color_count"] = ncol.value
        return out

    # --- differentiable rollout -------------------------------------------------
    def pin_count(self, i=0):
        return self.lib.pdx_pin_count(self.h, i)

    def set_pin_trajectory(self, i, traj):
        """traj: (nsub, pins, 3) pin targets per substep (pins in ascending vertex
        order), or None to clear."""
        if traj is None:
            self._check(self.lib.pdx_set_pin_trajectory(self.h, i, 0, None), "set_pin_trajectory")
            return
        t = np.ascontiguousarray(traj, np.float64)
        if t.ndim != 3 or t.shape[1] != self.pin_count(i) or t.shape[2] != 3:
            raise ValueError(f"set_pin_trajectory: expected (nsub, {self.pin_count(i)}, 3), got {t.shape}")
        self._check(self.lib.pdx_set_pin_trajectory(self.h, i, t.shape[0], _ptr(t)), "set_pin_trajectory")

    def rollout(self, nsub, checkpoint_every=10):
        self._check(self.lib.pdx_rollout(self.h, nsub, checkpoint_every), "rollout")
        self._roll_nsub = nsub

    def set_adjoint_seed(self, i, dx, dv=None):
        self._seeds = getattr(self, "_seeds", {})
        n3 = 3 * self._cloth(i).n
        dx = _arr(dx, np.float64, n3, "set_adjoint_seed dx")
        dv = None if dv is None else _arr(dv, np.float64, n3, "set_adjoint_seed dv")
        self._seeds[i] = (dx, dv)
        self._check(self.lib.pdx_set_adjoint_seed(self.h, i, _ptr(dx), None if dv is None else _ptr(dv)),
                    "set_adjoint_seed")

    def adjoint(self):
        self._check(self.lib.pdx_adjoint(self.h), "adjoint")

    def gradients(self, i=0):
        c = self.cloths[i]
        n, E, B, npin = c.n, len(c.edge_a), len(c.bend_a), self.pin_count(i)
        out = dict(x0=np.empty(3 * n), v0=np.empty(3 * n), kappa=np.empty(max(E, 1)),
                   bend=np.empty(max(B, 1)), pin=np.empty(max(self._roll_nsub * npin * 3, 1)))
        self._check(self.lib.pdx_get_gradients(self.h, i, _ptr(out["x0"]), _ptr(out["v0"]),
                                               _ptr(out["kappa"]), _ptr(out["bend"]), _ptr(out["pin"])),
                    "get_gradients")
        out["kappa"] = out["kappa"][:E]
        out["bend"] = out["bend"][:B]
        out["pin"] = out["pin"][:self._roll_nsub * npin * 3].reshape(self._roll_nsub, npin, 3)
        return out

    def rollout_timing(self):
        f, a = C.c_double(), C.c_double()
        self._check(self.lib.pdx_rollout_timing(self.h, C.byref(f), C.byref(a)), "rollout_timing")
        return f.value, a.value

    def set_schedule(self, block_threads, chain_cap=32):
        self._check(self.lib.pdx_set_schedule(self.h, block_threads, chain_cap), "set_schedule")

    def set_ordering(self, ordering):
        """Fill-reducing order of the factors: 'amd' (the reference's, default) or
        'nd' (nested dissection: fewer dependent sweep steps; fast mode only)."""
        code = {"amd": PDX_ORDER_AMD, "nd": PDX_ORDER_ND}[ordering]
        self._check(self.lib.pdx_set_ordering(self.h, code), "set_ordering")

    def set_split(self, min_factor_entries):
        """Instances whose factor has at least `min_factor_entries` entries run on two
        workgroups (mixed32, fast solve, nested-dissection order, batch launch); 0
        turns it off, 1 splits every instance that can be."""
        self._check(self.lib.pdx_set_split(self.h, int(min_factor_entries)), "set_split")

    def split_count(self):
        return self.lib.pdx_split_count(self.h)

    def split_threshold(self):
        return self.lib.pdx_get_split(self.h)

    def verify_schedule(self, i=0):
        """Host replay of the instance's sweep schedules in the kernel's phase order;
        raises PdxError on a dependency or bounds violation, else returns the max
        relative difference from a direct solve."""
        r = C.c_double()
        self._check(self.lib.pdx_verify_schedule(self.h, i, C.byref(r)), "verify_schedule")
        return r.value

    def set_launch(self, mode, grid_blocks=0):
        """mode: 'batch' (one workgroup per instance) or 'grid' (the whole GPU on
        one instance, fast solve only)."""
        m = {"batch": 0, "grid": 1}[mode]
        self._check(self.lib.pdx_set_launch(self.h, m, grid_blocks), "set_launch")

    def launch(self):
        """(launch form 'batch' / 'grid', grid workgroups (0 = all), fill order 'amd' / 'nd')."""
        m, g, o = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.pdx_get_launch(self.h, C.byref(m), C.byref(g), C.byref(o)), "get_launch")
        return ("batch", "grid")[m.value], g.value, ("amd", "nd")[o.value]

    def schedule_stats(self, i=0):
        o = np.zeros(16, np.int64)
        self._check(self.lib.pdx_get_schedule_stats(self.h, i, _ptr(o)), "get_schedule_stats")
        keys = ["steps", "tasks", "chain_tasks", "chain_rows", "runs", "dense", "max_tasks_per_step",
                "max_runs_per_step"]
        return {"fwd": dict(zip(keys, o[:8].tolist())), "bwd": dict(zip(keys, o[8:].tolist()))}

    def autotune(self, probe_substeps=1):
        """Picks the fastest launch shape for this context; returns (block, chain cap, ms)."""
        b, cap, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.pdx_autotune(self.h, probe_substeps, C.byref(b), C.byref(cap), C.byref(ms)),
                    "autotune")
        return b.value, cap.value, ms.value

    def last_kernel_ms(self):
        return self.lib.pdx_last_kernel_ms(self.h)

    def bytes_per_iteration(self):
        return self.lib.pdx_bytes_per_iteration(self.h)

    def bytes_per_substep(self):
        return self.lib.pdx_bytes_per_substep(self.h)

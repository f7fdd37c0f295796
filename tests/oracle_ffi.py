"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: the oracle restates the reference path on the CPU and is
the checker the GPU library is compared with.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_mesh.so")


class _Cfg(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("damping", C.c_double),
                ("max_iterations", C.c_int), ("rel_tol", C.c_double), ("abs_tol", C.c_double),
                ("residual_tol", C.c_double), ("cheb", C.c_int), ("cheb_rho", C.c_double),
                ("cheb_start", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, os.path.join(ORACLE_DIR, "liboracle.so")],
                           check=True, stdout=subprocess.DEVNULL)
        L = C.CDLL(ORACLE_LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, P, P, P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(_Cfg)]
        for name, res, args in [
            ("orc_destroy", None, [P]), ("orc_set_ordering", I, [P, P]), ("orc_set_threads", None, [P, I]),
            ("orc_step", I, [P, I]), ("orc_get_positions", None, [P, P]), ("orc_get_velocities", None, [P, P]),
            ("orc_set_positions", None, [P, P]), ("orc_set_velocities", None, [P, P]),
            ("orc_set_pin_positions", None, [P, P]), ("orc_set_edge_kappa", None, [P, P]),
            ("orc_set_config", None, [P, C.POINTER(_Cfg)]), ("orc_stats", None, [P, P]),
            ("orc_last_residual", D, [P]), ("orc_max_strain", D, [P]), ("orc_mean_strain", D, [P]),
            ("orc_surrogate_energy", D, [P]), ("orc_total_energy", D, [P]),
            ("orc_coloring", I, [P, P, P, P]), ("orc_factor", I, [P]),
            ("orc_compact_factor", I, [P, P, P, P, P, P, P]), ("orc_matrix", C.c_int64, [P, P, P, P]),
            ("orc_solve3", I, [P, P, P]), ("orc_nonlinear_residual", D, [P]), ("orc_bending_rhs", None, [P, P]),
            ("orc_set_bend_k", None, [P, P]), ("orc_set_masses", None, [P, P]),
            ("orc_set_pinned", None, [P, P]),
            ("orc_rollout_adjoint", I, [P, I, P, P, P, P, P, P, P, P]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _cfg(cfg):
    c = _Cfg()
    c.dt = cfg.dt
    for i in range(3):
        c.gravity[i] = cfg.gravity[i]
    c.damping = cfg.damping
    c.max_iterations = cfg.max_iterations
    c.rel_tol, c.abs_tol, c.residual_tol = cfg.rel_tol, cfg.abs_tol, cfg.residual_tol
    rho = getattr(cfg, "chebyshev_rho", 0.0)
    c.cheb, c.cheb_rho, c.cheb_start = int(rho > 0.0), rho, getattr(cfg, "chebyshev_start", 10)
    return c


class Oracle:
    """One cloth stepped by the FP64 CPU restatement of the reference."""

    def __init__(self, cloth, cfg, perm3, threads=1):
        L = lib()
        self.L = L
        self.cloth = cloth
        self._keep = cloth
        self.h = L.orc_create(cloth.n, _p(cloth.positions), _p(cloth.velocities), _p(cloth.masses),
                              _p(cloth.pinned), _p(cloth.pin_positions), cloth.n_edges, _p(cloth.edge_a),
                              _p(cloth.edge_b), _p(cloth.edge_rest), _p(cloth.edge_kappa), cloth.n_bends,
                              _p(cloth.bend_a) if cloth.n_bends else None,
                              _p(cloth.bend_b) if cloth.n_bends else None,
                              _p(cloth.bend_c) if cloth.n_bends else None,
                              _p(cloth.bend_k) if cloth.n_bends else None, C.byref(_cfg(cfg)))
        assert self.h
        perm3 = np.ascontiguousarray(perm3, np.int32)
        assert L.orc_set_ordering(self.h, _p(perm3)) == 0
        L.orc_set_threads(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def step(self, nsub=1):
        rc = self.L.orc_step(self.h, nsub)
        assert rc >= 0, rc
        return rc

    def factor(self):
        assert self.L.orc_factor(self.h) == 0

    def positions(self):
        x = np.empty(3 * self.cloth.n)
        self.L.orc_get_positions(self.h, _p(x))
        return x

    def velocities(self):
        v = np.empty(3 * self.cloth.n)
        self.L.orc_get_velocities(self.h, _p(v))
        return v

    def set_pin_positions(self, q):
        q = np.ascontiguousarray(q, np.float64)
        self.L.orc_set_pin_positions(self.h, _p(q))

    def set_edge_kappa(self, k):
        k = np.ascontiguousarray(k, np.float64)
        self.L.orc_set_edge_kappa(self.h, _p(k))

    def set_config(self, cfg):
        self.L.orc_set_config(self.h, C.byref(_cfg(cfg)))

    def set_bend_k(self, k):
        self.L.orc_set_bend_k(self.h, _p(np.ascontiguousarray(k, np.float64)))

    def set_masses(self, m):
        self.L.orc_set_masses(self.h, _p(np.ascontiguousarray(m, np.float64)))

    def set_pinned(self, p):
        self.L.orc_set_pinned(self.h, _p(np.ascontiguousarray(np.asarray(p) != 0, np.uint8)))

    def set_ordering(self, perm3):
        perm3 = np.ascontiguousarray(perm3, np.int32)
        assert self.L.orc_set_ordering(self.h, _p(perm3)) == 0

    def rollout_adjoint(self, nsub, pin_traj, seed_x, seed_v=None):
        """FP64 CPU adjoint of nsub fixed-iteration substeps (advances the state).
        Returns dict of gradients: x0, v0, kappa (per edge), bend (per stencil),
        pin (nsub x pins x 3)."""
        n, E, B = self.cloth.n, self.cloth.n_edges, len(self.cloth.bend_a)
        npin = int(self.cloth.pinned.sum())
        out = dict(x0=np.zeros(3 * n), v0=np.zeros(3 * n), kappa=np.zeros(E), bend=np.zeros(max(B, 1)),
                   pin=np.zeros(nsub * npin * 3))
        traj = None if pin_traj is None else np.ascontiguousarray(pin_traj, np.float64).reshape(-1)
        sx = np.ascontiguousarray(seed_x, np.float64)
        sv = None if seed_v is None else np.ascontiguousarray(seed_v, np.float64)
        rc = self.L.orc_rollout_adjoint(self.h, nsub, None if traj is None else _p(traj), _p(sx),
                                        None if sv is None else _p(sv), _p(out["x0"]), _p(out["v0"]),
                                        _p(out["kappa"]), _p(out["bend"]), _p(out["pin"]))
        if rc != 0:
            raise RuntimeError(f"orc_rollout_adjoint failed: {rc}")
        out["bend"] = out["bend"][:B]
        out["pin"] = out["pin"].reshape(nsub, npin, 3)
        return out

    def stats(self):
        s = np.zeros(9, np.int64)
        self.L.orc_stats(self.h, _p(s))
        keys = ["iterations_used", "converged", "early_exit_count", "factorize_count", "analyze_count",
                "color_count", "compact_nnz", "components", "compact_ready"]
        return dict(zip(keys, s.tolist()))

    def last_residual(self):
        return self.L.orc_last_residual(self.h)

    def max_strain(self):
        return self.L.orc_max_strain(self.h)

    def mean_strain(self):
        return self.L.orc_mean_strain(self.h)

    def surrogate_energy(self):
        return self.L.orc_surrogate_energy(self.h)

    def total_energy(self):
        return self.L.orc_total_energy(self.h)

    def coloring(self):
        E = self.cloth.n_edges
        col, order = np.empty(E, np.uint32), np.empty(E, np.uint32)
        start = np.empty(E + 2, np.uint32)
        cc = self.L.orc_coloring(self.h, _p(col), _p(order), _p(start))
        return col, order, start[:cc + 1]

    def compact_factor(self):
        self.factor()
        nnz = np.zeros(1, np.int64)
        m = self.L.orc_compact_factor(self.h, None, None, None, None, None, _p(nnz))
        assert m > 0
        out = dict(ptr=np.empty(m + 1, np.int32), rows=np.empty(nnz[0], np.int32), vals=np.empty(nnz[0]),
                   dinv=np.empty(m), comp_dof=np.empty(3 * m, np.int32))
        self.L.orc_compact_factor(self.h, _p(out["ptr"]), _p(out["rows"]), _p(out["vals"]), _p(out["dinv"]),
                                  _p(out["comp_dof"]), _p(nnz))
        return out

    def matrix(self):
        self.factor()
        nnz = self.L.orc_matrix(self.h, None, None, None)
        cp = np.empty(self.cloth.n + 1, np.int32)
        r = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        self.L.orc_matrix(self.h, _p(cp), _p(r), _p(v))
        return cp, r, v

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        assert self.L.orc_solve3(self.h, _p(b), _p(x)) == 0
        return x

    def nonlinear_residual(self):
        return self.L.orc_nonlinear_residual(self.h)


def ref_mesh_lib():
    """The reference's own Mesh/ConstraintColoring sources built into oracle/_ref
    (only where the reference checkout was present at build time)."""
    if not os.path.exists(REF_LIB):
        return None
    L = C.CDLL(REF_LIB)
    P = C.c_void_p
    L.ref_grid_scene.restype = P
    L.ref_grid_scene.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double]
    for name, args in [("ref_free", [P]), ("ref_counts", [P, P]), ("ref_vertices", [P, P, P, P]),
                       ("ref_edges", [P, P, P, P, P]), ("ref_bends", [P, P, P, P, P]),
                       ("ref_coloring", [P, P, P, P, P, P]), ("ref_bend_adjacency", [P, P, P, P])]:
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = args
    if hasattr(L, "ref_load_obj"):
        L.ref_load_obj.restype = P
        L.ref_load_obj.argtypes = [C.c_char_p, C.c_double, C.c_double]
        L.ref_triangle_count.restype = C.c_int64
        L.ref_triangle_count.argtypes = [P]
        L.ref_triangles.restype = None
        L.ref_triangles.argtypes = [P, P]
    return L

"""The FP64 CPU adjoint of the oracle against central finite differences of the
oracle's own forward substeps (step 1e-6, FP64).  CPU only.

Scene: a small grid with distance, shear and bend constraints, pins on the top
edge (one in its middle, so that bend stencils and edges straddle pinned and free
vertices), fixed iterations, a moving pin trajectory, seeded v0.  Loss:
sum |x_T - target|^2.  Every gradient the adjoint returns (x0, v0, per-edge kappa,
per-stencil k, pin trajectory) is compared as a directional derivative along a
seeded random direction, plus a few single entries."""
import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes
from oracle_ffi import Oracle

EPS = 1e-6      # the step of the check: central differences in FP64
RTOL = 1e-4     # at that step the differences carry ~1e-6 relative rounding noise
EPS_T = 1e-4    # a larger step with Richardson extrapolation (truncation O(eps^4))
RTOL_T = 1e-7   # pins the adjoint well below the rounding noise of the 1e-6 step


def _scene(nx=9, ny=7, h=1.0 / 240.0, iters=6):
    top = (ny - 1) * nx
    c = scenes.grid_cloth(nx, ny, 800.0, shear=160.0, bend=16.0, pins=[top, top + nx // 2, top + nx - 1])
    rng = np.random.default_rng(5)
    c.velocities = 0.1 * rng.standard_normal(3 * c.n)
    cfg = SimConfig.fixed_iterations(iters, h)
    e = Engine("fp64")
    e.add(c, cfg)
    perm = e.factor_arrays(0)["perm3"]
    e.close()
    return c, cfg, perm


def _traj(c, nsub, h):
    pins = np.flatnonzero(c.pinned)
    q0 = c.positions.reshape(-1, 3)[pins]
    t = (np.arange(nsub) + 1) * h
    off = 0.05 * np.sin(2 * np.pi * 2.0 * t)
    traj = np.repeat(q0[None], nsub, axis=0).copy()
    traj[:, :, 1] += off[:, None]          # vertical sinusoid of the pins (2 Hz, 0.05 m)
    traj[:, 1, 2] += 0.5 * off             # the middle pin also moves along z
    return traj


class Rollout:
    def __init__(self, c, cfg, perm, nsub, target):
        self.c, self.cfg, self.perm, self.nsub, self.target = c, cfg, perm, nsub, target
        self.pins = np.flatnonzero(c.pinned)

    def run(self, x0=None, v0=None, kappa=None, bend=None, traj=None):
        c = self.c
        o = Oracle(c, self.cfg, self.perm)
        if x0 is not None:
            o.L.orc_set_positions(o.h, np.ascontiguousarray(x0).ctypes.data)
        if v0 is not None:
            o.L.orc_set_velocities(o.h, np.ascontiguousarray(v0).ctypes.data)
        if kappa is not None:
            o.set_edge_kappa(kappa)
        if bend is not None:
            o.set_bend_k(bend)
        q = c.pin_positions.copy().reshape(-1, 3)
        for s in range(self.nsub):
            q[self.pins] = traj[s]
            o.set_pin_positions(q.reshape(-1))
            assert o.step(1) == self.cfg.max_iterations
        x = o.positions()
        return float(np.sum((x - self.target) ** 2)), x

    def adjoint(self, traj):
        _, xT = self.run(traj=traj)
        o = Oracle(self.c, self.cfg, self.perm)
        g = o.rollout_adjoint(self.nsub, traj, 2.0 * (xT - self.target))
        assert np.array_equal(o.positions(), xT)  # the adjoint's forward is the same forward
        return g


@pytest.fixture(scope="module")
def setup():
    c, cfg, perm = _scene()
    nsub = 5
    traj = _traj(c, nsub, cfg.dt)
    r0 = Rollout(c, cfg, perm, nsub, np.zeros(3 * c.n))
    _, xT = r0.run(traj=traj)
    target = xT + 0.01 * np.random.default_rng(9).standard_normal(xT.shape)
    r = Rollout(c, cfg, perm, nsub, target)
    return r, traj, r.adjoint(traj)


def _fd(f, base, d, eps=EPS):
    lp, _ = f(base + eps * d)
    lm, _ = f(base - eps * d)
    return (lp - lm) / (2 * eps)


def _check(fd, ad, rtol=RTOL):
    assert abs(fd - ad) <= rtol * max(abs(fd), abs(ad)) + 1e-14, (fd, ad)


def _check_both(f, base, d, ad):
    _check(_fd(f, base, d), ad)
    rich = (4.0 * _fd(f, base, d, EPS_T / 2) - _fd(f, base, d, EPS_T)) / 3.0
    _check(rich, ad, RTOL_T)


def test_grad_v0(setup):
    r, traj, g = setup
    c = r.c
    rng = np.random.default_rng(1)
    d = rng.standard_normal(3 * c.n)
    d[np.repeat(c.pinned.astype(bool), 3)] = 0.0
    _check_both(lambda v: r.run(v0=v, traj=traj), c.velocities, d, g["v0"] @ d)
    for i in (0, 7, 3 * (c.n // 2) + 1):
        e = np.zeros(3 * c.n)
        e[i] = 1.0
        _check(_fd(lambda v: r.run(v0=v, traj=traj), c.velocities, e), g["v0"][i])


def test_grad_x0(setup):
    r, traj, g = setup
    c = r.c
    d = np.random.default_rng(2).standard_normal(3 * c.n)  # pinned entries included
    _check_both(lambda x: r.run(x0=x, traj=traj), c.positions, d, g["x0"] @ d)


def test_grad_kappa(setup):
    r, traj, g = setup
    c = r.c
    d = np.random.default_rng(3).standard_normal(c.n_edges) * c.edge_kappa
    _check_both(lambda k: r.run(kappa=k, traj=traj), c.edge_kappa, d, g["kappa"] @ d)
    # instance stiffness scale: every edge and stencil scaled together
    _check_both(lambda s: r.run(kappa=c.edge_kappa * s[0], bend=c.bend_k * s[0], traj=traj),
                np.ones(1), np.ones(1), g["kappa"] @ c.edge_kappa + g["bend"] @ c.bend_k)


def test_grad_bend(setup):
    r, traj, g = setup
    c = r.c
    d = np.random.default_rng(4).standard_normal(len(c.bend_k)) * c.bend_k
    _check_both(lambda k: r.run(bend=k, traj=traj), c.bend_k, d, g["bend"] @ d)


def test_grad_pin_trajectory(setup):
    r, traj, g = setup
    d = np.random.default_rng(6).standard_normal(traj.shape)
    _check_both(lambda t: r.run(traj=t), traj, d, float(np.sum(g["pin"] * d)))
    e = np.zeros(traj.shape)
    e[-1, 1, 1] = 1.0
    _check(_fd(lambda t: r.run(traj=t), traj, e), g["pin"][-1, 1, 1])
    e = np.zeros(traj.shape)
    e[0, 0, 2] = 1.0
    _check(_fd(lambda t: r.run(traj=t), traj, e), g["pin"][0, 0, 2])

"""The workloads at their full sizes on the GPU, checked against the FP64 paths:

* config 1 (40 x 40, 600 substeps) and config 2 (100 x 100 with shear and bend,
  300 substeps) in every FP64 form and the FP32-solve form against the CPU
  oracle over the whole run;
* the 512-instance mix, 200 substeps, through pdx_step_graph (the benchmark's
  path): the FP32-solve form in the nested-dissection order against the FP64
  batch in the reference's AMD order, every instance within 1e-5 m, and four
  sampled instances against the FP64 CPU oracle;
* the 512 x 512 cloth: the FP64 grid form against the CPU oracle after one
  substep, the FP32-solve grid form against the FP64 one after three;
* the differentiable batch, two 128 x 128 instances, 120 substeps, 20 iterations,
  checkpoints every 10: GPU adjoint gradients (FP64 and FP32 solves) against
  central differences of the GPU FP64 forward, and one instance against the FP64
  CPU adjoint of the oracle.

The CPU oracle runs on host threads while the GPU works.  Every GPU step prints
its case first."""
import threading
import time

import numpy as np
import pytest

from pd_ipc_amd import Cloth, Engine, SimConfig, scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

T0 = time.time()


def _say(*a):
    print(f"[case {time.time() - T0:7.1f}s]", *a, flush=True)


class _Bg(threading.Thread):
    """Runs fn() on a host thread (the oracle's ctypes calls release the GIL)."""

    def __init__(self, fn):
        super().__init__(daemon=True)
        self.fn, self.out, self.err = fn, None, None
        self.start()

    def run(self):
        try:
            self.out = self.fn()
        except BaseException as e:  # re-raised by result()
            self.err = e

    def result(self):
        self.join()
        if self.err is not None:
            raise self.err
        return self.out


def _amd_perm(c, cfg):
    e = Engine("fp64", "exact")
    e.add(c, cfg)
    perm = e.factor_arrays(0)["perm3"]
    e.close()
    return perm


def _run(c, cfg, nsub, precision, solve, ordering=None, launch=None):
    e = Engine(precision, solve)
    e.add(c, cfg)
    if ordering:
        e.set_ordering(ordering)
    if launch:
        e.set_launch(*launch)
    e.step(nsub)
    x = e.state(0)[0]
    st = e.stats(0)
    e.close()
    assert st["factorize_count"] == 1 and st["substeps"] == nsub
    return x


@pytest.mark.timeout(900)
def test_config1_full_length():
    """40 x 40, distance only, 10 iterations, 600 substeps, against the FP64 CPU
    oracle (AMD order) over the whole run: the exact FP64 form bit for bit, the
    dense and the sparse (nested dissection, one workgroup) FP64 forms within
    1e-10 relative, the FP32-solve form within 1e-5 m."""
    c, cfg, nsub = scenes.config1()
    perm = _amd_perm(c, cfg)
    orc = _Bg(lambda: _oracle_positions(c, cfg, perm, nsub))
    _say("config 1: exact FP64,", nsub, "substeps")
    xe = _run(c, cfg, nsub, "fp64", "exact")
    _say("config 1: dense FP64")
    xd = _run(c, cfg, nsub, "fp64", "dense")
    _say("config 1: sparse FP64, ND")
    xs = _run(c, cfg, nsub, "fp64", "fast", "nd")
    _say("config 1: sparse mixed32, ND")
    x32 = _run(c, cfg, nsub, "mixed32", "fast", "nd")
    xo = orc.result()
    scale = np.abs(xo).max()
    assert np.array_equal(xe.view(np.uint64), xo.view(np.uint64))
    for x in (xd, xs):
        assert np.abs(x - xo).max() <= 1e-10 * scale
    assert np.abs(x32 - xo).max() <= 1e-5


@pytest.mark.timeout(900)
def test_config2_full_length():
    """100 x 100 with shear and bending, sigma 1e-3 m perturbation, 40 iterations,
    300 substeps, against the FP64 CPU oracle over the whole run: the dense FP64
    form and the grid form (level sweeps, nested dissection) in FP64 within 1e-10
    relative, the FP32-solve grid form within 1e-5 m; the exact FP64 form bit for
    bit over its first 2 substeps (1.5 s each on the GPU: the reference's own
    dependent chain of subtractions)."""
    c, cfg, nsub = scenes.config2()
    perm = _amd_perm(c, cfg)
    orc = _Bg(lambda: _oracle_positions(c, cfg, perm, nsub))
    orc2 = _oracle_positions(c, cfg, perm, 2)
    _say("config 2: exact FP64, 2 substeps")
    xe = _run(c, cfg, 2, "fp64", "exact")
    assert np.array_equal(xe.view(np.uint64), orc2.view(np.uint64))
    _say("config 2: dense FP64,", nsub, "substeps")
    xd = _run(c, cfg, nsub, "fp64", "dense")
    _say("config 2: grid form FP64, ND")
    xg = _run(c, cfg, nsub, "fp64", "fast", "nd", ("grid", 32))
    _say("config 2: grid form mixed32, ND")
    x32 = _run(c, cfg, nsub, "mixed32", "fast", "nd", ("grid", 32))
    _say("config 2: waiting for the oracle")
    xo = orc.result()
    scale = np.abs(xo).max()
    for x in (xd, xg):
        assert np.abs(x - xo).max() <= 1e-10 * scale
    assert np.abs(x32 - xo).max() <= 1e-5


@pytest.mark.timeout(900)
def test_mix_512_full_size_through_graph():
    cloths, cfg, nsub = scenes.config4()
    assert len(cloths) == 512 and nsub == 200
    sizes = np.array([c.n for c in cloths])
    bends = [i for i, c in enumerate(cloths) if c.n_bends > 0]
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), bends[0], bends[len(bends) // 2]})
    _say("mix: factor FP64 AMD and mixed32 ND contexts,", int(sizes.sum()), "vertices")
    e64 = Engine("fp64", "fast")
    e32 = Engine("mixed32", "fast")
    for c in cloths:
        e64.add(c, cfg)
        e32.add(c, cfg)
    e32.set_ordering("nd")
    e64.factor()
    e32.factor()
    # the oracle on the sampled instances (reference AMD order), on host threads
    perms = {i: e64.factor_arrays(i)["perm3"] for i in sample}

    def oracle_run(i):
        o = Oracle(cloths[i], cfg, perms[i], threads=4)
        o.step(nsub)
        return o.positions()

    bg = {i: _Bg(lambda i=i: oracle_run(i)) for i in sample}
    _say("mix: mixed32 ND, 200 substeps as 200 graph replays")
    e32.step_graph(1, nsub)
    _say("mix: FP64 AMD, 200 substeps")
    e64.step_graph(1, nsub)
    err = np.zeros(len(cloths))
    for i in range(len(cloths)):
        x32, _ = e32.state(i)
        x64, _ = e64.state(i)
        assert np.all(np.isfinite(x32))
        err[i] = np.abs(x32 - x64).max()
        s = e32.stats(i)
        assert s["substeps"] == nsub and s["factorize_count"] == 1 and s["iterations_used"] == 40
    print("max |x32 - x64| over instances", err.max(), "at", int(err.argmax()), "median", np.median(err))
    assert err.max() <= 1e-5
    for i in sample:
        _say("mix: oracle instance", i, "n =", cloths[i].n)
        xo = bg[i].result()
        x64, _ = e64.state(i)
        x32, _ = e32.state(i)
        rel = np.abs(x64 - xo).max() / np.abs(xo).max()
        e32o = np.abs(x32 - xo).max()
        print("instance", i, "FP64 batch vs oracle rel", rel, "mixed32 vs oracle", e32o)
        assert rel <= 1e-10
        assert e32o <= 1e-5


@pytest.mark.timeout(900)
def test_large_cloth_512_grid_form():
    """512 x 512 with shear and bending, four corners pinned, 40 iterations: the
    grid form in FP64 (nested-dissection order) against the FP64 CPU oracle (the
    reference path restated, AMD order) after one substep, within 1e-10
    relative; the FP32-solve grid form against the FP64 grid form after 3
    substeps, within 1e-5 m; one factorisation each."""
    c, cfg, _ = scenes.config3()
    nsub = 3
    _say("512x512: AMD order for the oracle")
    amd = Engine("fp64", "exact")
    amd.add(c, cfg)
    perm = amd.factor_arrays(0)["perm3"]
    amd.close()
    orc = _Bg(lambda: _oracle_positions(c, cfg, perm, 1))
    engines = {}
    for prec, g in (("fp64", 128), ("mixed32", 64)):
        _say("512x512: factor", prec, "ND grid form,", g, "workgroups")
        e = Engine(prec, "fast")
        e.add(c, cfg)
        e.set_ordering("nd")
        e.set_launch("grid", g)
        e.set_schedule(256, 32)
        e.factor()
        engines[prec] = e
    e64, e32 = engines["fp64"], engines["mixed32"]
    _say("512x512: FP64 grid form, 1 substep")
    e64.step(1)
    x1 = e64.state(0)[0].copy()
    _say("512x512: FP64 grid form, 2 more substeps; mixed32 grid form,", nsub, "substeps")
    e64.step(nsub - 1)
    e32.step(nsub)
    x64, x32 = e64.state(0)[0], e32.state(0)[0]
    err = np.abs(x32 - x64).max()
    print("512x512 max |x32 - x64|", err, "after", nsub, "substeps")
    assert np.all(np.isfinite(x32)) and err <= 1e-5
    for e in (e32, e64):
        s = e.stats(0)
        assert s["factorize_count"] == 1 and s["substeps"] == nsub and s["iterations_used"] == 40
        e.close()
    _say("512x512: waiting for the oracle (factor and 1 substep)")
    orc.join()
    assert orc.err is None, orc.err
    xo = orc.out
    rel = np.abs(x1 - xo).max() / np.abs(xo).max()
    print("512x512 FP64 grid form vs oracle after 1 substep, relative", rel)
    assert rel <= 1e-10


def _oracle_positions(c, cfg, perm, nsub):
    o = Oracle(c, cfg, perm, threads=4)
    o.step(nsub)
    return o.positions()


def _scaled(c, s):
    d = Cloth(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.__dict__.items()})
    d.edge_kappa = c.edge_kappa * s
    d.bend_k = c.bend_k * s
    return d


def _with_v0(c, v0):
    d = Cloth(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.__dict__.items()})
    d.velocities = v0
    return d


def _rollout(cloths, trajs, cfg, nsub, precision, every=10, seeds=None):
    e = Engine(precision, "fast")
    for c in cloths:
        e.add(c, cfg)
    e.set_ordering("nd")
    for i, t in enumerate(trajs):
        e.set_pin_trajectory(i, t)
    e.rollout(nsub, every)
    xs = [e.state(i)[0] for i in range(len(cloths))]
    g = None
    if seeds is not None:
        for i, s in enumerate(seeds):
            e.set_adjoint_seed(i, s(xs[i]))
        e.adjoint()
        g = [e.gradients(i) for i in range(len(cloths))]
        fwd, adj = e.rollout_timing()
        print(precision, "rollout", round(fwd, 1), "ms, adjoint", round(adj, 1), "ms")
    e.close()
    return xs, g


@pytest.mark.timeout(900)
def test_adjoint_config5_two_instances_full_length():
    cloths, cfg, nsub, hidden = scenes.config5(count=2)
    assert cloths[0].n == 128 * 128 and nsub == 120 and cfg.max_iterations == 20
    base_traj = [scenes.pin_trajectory(c, nsub, cfg.dt, amplitude=0.0) for c in cloths]  # static pins
    _say("config 5: hidden forward (FP64) for the targets")
    targets, _ = _rollout([_scaled(c, h["kappa"] / 2300.0) for c, h in zip(cloths, hidden)],
                          [h["traj"] for h in hidden], cfg, nsub, "fp64")
    seeds = [lambda x, t=t: 2.0 * (x - t) for t in targets]
    # FP64 CPU adjoint of instance 0 (reference AMD order) on host threads
    e = Engine("fp64")
    e.add(cloths[0], cfg)
    perm = e.factor_arrays(0)["perm3"]
    e.close()

    def cpu_adjoint():
        o = Oracle(cloths[0], cfg, perm, threads=8)
        o_fwd = Oracle(cloths[0], cfg, perm, threads=8)
        q = cloths[0].pin_positions.copy().reshape(-1, 3)
        pins = np.flatnonzero(cloths[0].pinned)
        for s in range(nsub):  # the forward of the loss, with the same pin targets
            q[pins] = base_traj[0][s]
            o_fwd.set_pin_positions(q.reshape(-1))
            o_fwd.step(1)
        xT = o_fwd.positions()
        return xT, o.rollout_adjoint(nsub, base_traj[0], 2.0 * (xT - targets[0]))

    bg = _Bg(cpu_adjoint)
    _say("config 5: GPU adjoint, FP64 solves")
    x64, g64 = _rollout(cloths, base_traj, cfg, nsub, "fp64", seeds=seeds)
    _say("config 5: GPU adjoint, FP32 solves")
    x32, g32 = _rollout(cloths, base_traj, cfg, nsub, "mixed32", seeds=seeds)
    print("max |x32 - x64| after the rollout", max(np.abs(a - b).max() for a, b in zip(x32, x64)))
    # central differences of the GPU FP64 forward, all perturbed copies in one launch:
    # per instance, stiffness scale at a relative step of 1e-3 (the FP64 forward's
    # ~1e-12 m rounding noise swamps a 1e-6 step on this parameter), velocities and
    # pin trajectory at a step of 1e-6 along seeded directions
    eps, ks = 1e-6, 1e-3
    rng = np.random.default_rng(55)
    dirs_v, dirs_p, fd_cloths, fd_trajs = [], [], [], []
    for c, t in zip(cloths, base_traj):
        dv = rng.standard_normal(3 * c.n)
        dv[np.repeat(c.pinned.astype(bool), 3)] = 0.0
        dp = rng.standard_normal(t.shape)
        dirs_v.append(dv)
        dirs_p.append(dp)
        fd_cloths += [_scaled(c, 1 + ks), _scaled(c, 1 - ks), _with_v0(c, c.velocities + eps * dv),
                      _with_v0(c, c.velocities - eps * dv), c, c]
        fd_trajs += [t, t, t, t, t + eps * dp, t - eps * dp]
    _say("config 5: central differences, 12 perturbed rollouts in one launch (FP64)")
    xs, _ = _rollout(fd_cloths, fd_trajs, cfg, nsub, "fp64")
    for i in range(2):
        L = [float(np.sum((x - targets[i]) ** 2)) for x in xs[6 * i:6 * i + 6]]
        fd = {"kappa": (L[0] - L[1]) / (2 * ks), "v0": (L[2] - L[3]) / (2 * eps), "pin": (L[4] - L[5]) / (2 * eps)}
        for name, g in (("fp64", g64[i]), ("mixed32", g32[i])):
            ad = {"kappa": g["kappa"] @ cloths[i].edge_kappa + g["bend"] @ cloths[i].bend_k,
                  "v0": g["v0"] @ dirs_v[i], "pin": float(np.sum(g["pin"] * dirs_p[i]))}
            for k in fd:
                rel = abs(ad[k] - fd[k]) / abs(fd[k])
                print(f"instance {i} {name} {k}: fd {fd[k]:.10g} adjoint {ad[k]:.10g} rel {rel:.2e}")
                assert rel <= 1e-4, (i, name, k)
    _say("config 5: FP64 CPU adjoint of instance 0")
    xT_cpu, gc = bg.result()
    print("GPU FP64 (ND order) vs CPU oracle (AMD order) xT", np.abs(x64[0] - xT_cpu).max())
    for k in ("x0", "v0", "kappa", "bend", "pin"):
        a, b = np.asarray(g64[0][k]).reshape(-1), np.asarray(gc[k]).reshape(-1)
        rel = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        a32 = np.asarray(g32[0][k]).reshape(-1)
        rel32 = np.abs(a32 - b).max() / max(np.abs(b).max(), 1e-300)
        print(f"{k}: GPU FP64 vs CPU adjoint {rel:.2e}, GPU mixed32 vs CPU adjoint {rel32:.2e}")
        assert rel <= 1e-7, k

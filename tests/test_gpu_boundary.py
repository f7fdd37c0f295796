"""The C-ABI boundary on the GPU: launch entry points that must leave the same
bits (pdx_step, pdx_step_async + pdx_sync, pdx_step_graph), the reference's
factor-once rules replayed step by step against the oracle
(tests/chain/spring_vertical.cpp:312-378 and :528-563), contexts driven from two
host threads, instances added after stepping, and launch-shape changes of the
grid form.  Every GPU step prints its case first."""
import threading

import numpy as np
import pytest

from pd_ipc_amd import Cloth, Engine, SimConfig, scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _same(a, b):
    return np.array_equal(_bits(a), _bits(b))


def _mix(count, seed=41):
    """Small uneven cloths of the batch mix (sides 16..40 so the test stays short)."""
    out = []
    for i in range(count):
        c = scenes.config4_instance(seed, i)
        nx, ny = 16 + (c.meta["index"] * 7) % 25, 16 + (c.meta["index"] * 11) % 25
        k = c.meta["kappa"]
        d = scenes.grid_cloth(nx, ny, k, shear=c.meta["shear"], bend=c.meta["bend"],
                              pins=[(ny - 1) * nx, (ny - 1) * nx + nx - 1])
        d.velocities = 0.1 * np.random.default_rng(seed + i).standard_normal(3 * d.n)
        d.velocities.reshape(-1, 3)[d.pinned != 0] = 0.0
        out.append(d)
    return out


def _engine(cloths, cfg, precision, solve, ordering="amd"):
    e = Engine(precision, solve)
    for c in cloths:
        e.add(c, cfg)
    if ordering != "amd":
        e.set_ordering(ordering)
    e.factor()
    return e


@pytest.mark.parametrize("precision,solve,ordering", [("mixed32", "fast", "nd"), ("fp64", "exact", "amd"),
                                                      ("fp64", "fast", "nd")])
def test_step_graph_and_async_match_step(precision, solve, ordering):
    """pdx_step_graph(1, K) (the bench's timed path) and pdx_step_async + pdx_sync
    leave the state of pdx_step(K) bit for bit, and advance the counters alike."""
    K = 7
    cloths = _mix(6)
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    _say("step vs graph vs async", precision, solve, ordering, K)
    a = _engine(cloths, cfg, precision, solve, ordering)
    b = _engine(cloths, cfg, precision, solve, ordering)
    c = _engine(cloths, cfg, precision, solve, ordering)
    a.step(K)
    b.step_graph(1, K)
    c.step_async(K)
    c.sync()
    assert b.last_kernel_ms() > 0.0
    for i in range(len(cloths)):
        xa, va = a.state(i)
        for other in (b, c):
            x, v = other.state(i)
            assert _same(x, xa) and _same(v, va), i
            s = other.stats(i)
            assert s["substeps"] == K and s["total_iterations"] == 12 * K and s["factorize_count"] == 1
    # a second graph call continues from where the first left the state
    a.step(3)
    b.step_graph(1, 3)
    for i in range(len(cloths)):
        assert _same(b.state(i)[0], a.state(i)[0])
        assert b.stats(i)["substeps"] == K + 3


def test_grid_form_replays_directly():
    """The cooperative grid form is replayed launch by launch by pdx_step_graph:
    same bits as pdx_step."""
    c = scenes.grid_cloth(48, 40, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    _say("grid form graph replay")
    engines = []
    for _ in range(2):
        e = Engine("mixed32", "fast")
        e.add(c, cfg)
        e.set_ordering("nd")
        e.set_launch("grid", 32)
        e.factor()
        engines.append(e)
    engines[0].step(4)
    engines[1].step_graph(1, 4)
    assert _same(engines[0].state(0)[0], engines[1].state(0)[0])


def _spring(y0=0.12, mass=0.05, rest=0.1, k=1.0e4):
    """makeTwoVertexSpring (tests/chain/spring_vertical.cpp:27-53): vertex 0 pinned
    at the origin, vertex 1 at height y0, one edge."""
    pos = np.array([0.0, 0.0, 0.0, 0.0, y0, 0.0])
    return Cloth(positions=pos, masses=np.array([mass, mass]), pinned=np.array([1, 0], np.uint8),
                 edge_a=np.array([0], np.int32), edge_b=np.array([1], np.int32), edge_rest=np.array([rest]),
                 edge_kappa=np.array([k]))


def test_spring_vertical_factor_once_replay():
    """globalMatrixIsFactoredOnceAndReused (spring_vertical.cpp:312-378) on the GPU,
    exact FP64, in lockstep with the oracle: one factorisation over 2001 substeps;
    kappa x2 -> 2; h/2 -> 3; pin mask -> 4 (and a new analysis); moving a pin and
    changing damping over 100 substeps -> still 4.  Positions are the oracle's bit
    for bit after every stage."""
    h = 1.0 / 120.0
    c = _spring()
    cfg = SimConfig(dt=h, damping=0.0, max_iterations=40, rel_tol=0.0, abs_tol=1e-16, residual_tol=1e-3)
    e = Engine("fp64", "exact")
    e.add(c, cfg)
    e.factor()
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])

    def stage(nsub, fc, ac, what):
        _say("spring", what, nsub)
        e.step(nsub)
        o.step(nsub)
        x, v = e.state(0)
        assert _same(x, o.positions()) and _same(v, o.velocities()), what
        s, so = e.stats(0), o.stats()
        assert (s["factorize_count"], s["analyze_count"]) == (fc, ac), (what, s)
        assert (so["factorize_count"], so["analyze_count"]) == (fc, ac), (what, so)
        return x

    stage(1, 1, 1, "first step")
    x = stage(2000, 1, 1, "2000 steps")
    # the equilibrium rest - m g / kappa (spring_vertical.cpp:353)
    assert abs(x[4] - (0.1 - 0.05 * 9.81 / 1.0e4)) / 0.1 <= 1e-6
    e.set_edge_kappa(0, [2.0e4])
    o.set_edge_kappa(np.array([2.0e4]))
    stage(1, 2, 1, "kappa x2")
    cfg2 = SimConfig(dt=h * 0.5, damping=0.0, max_iterations=40, rel_tol=0.0, abs_tol=1e-16, residual_tol=1e-3)
    e.set_config(0, cfg2)
    o.set_config(cfg2)
    x = stage(1, 3, 1, "h / 2")
    q = c.pin_positions.copy()
    q[3:6] = x[3:6]  # pin vertex 1 where it is
    e.set_pinned(0, [1, 1])
    e.set_pin_positions(0, q)
    o.set_pinned(np.array([1, 1]))
    o.set_pin_positions(q)
    o.set_ordering(e.factor_arrays(0)["perm3"])
    stage(1, 4, 2, "pin mask")
    q[0:3] = [1.0, -0.5, 0.25]
    e.set_pin_positions(0, q)
    o.set_pin_positions(q)
    cfg3 = SimConfig(dt=h * 0.5, damping=0.3, max_iterations=40, rel_tol=0.0, abs_tol=1e-16, residual_tol=1e-3)
    e.set_config(0, cfg3)
    o.set_config(cfg3)
    x = stage(100, 4, 2, "move pin + damping")
    assert np.all(np.isfinite(x)) and np.array_equal(x[0:3], q[0:3])


def test_pin_mask_change_reanalyses_replay():
    """pinMaskChangeReanalyzesSparsityPattern (spring_vertical.cpp:528-563): a 6x6
    makeScene cloth (spacing 0.05, kappa 1e4, edge-lumped masses, 4 iterations,
    rel = abs = 0); pinning interior vertex 14 analyses once more, then neither
    analysis nor factorisation repeats.  Exact FP64 against the oracle, bitwise."""
    c = scenes.grid_cloth(6, 6, 1.0e4, sx=0.05, sz=0.05, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=4, rel_tol=0.0, abs_tol=0.0, residual_tol=1e-3)
    e = Engine("fp64", "exact")
    e.add(c, cfg)
    e.factor()
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    _say("pin mask 6x6")
    e.step(1)
    o.step(1)
    assert e.stats(0)["analyze_count"] == 1
    e.step(5)
    o.step(5)
    assert e.stats(0)["analyze_count"] == 1
    v = 2 * 6 + 2
    assert c.pinned[v] == 0
    x, _ = e.state(0)
    pinned = c.pinned.copy()
    pinned[v] = 1
    q = c.pin_positions.copy()
    q[3 * v:3 * v + 3] = x[3 * v:3 * v + 3]
    e.set_pinned(0, pinned)
    e.set_pin_positions(0, q)
    o.set_pinned(pinned)
    o.set_pin_positions(q)
    o.set_ordering(e.factor_arrays(0)["perm3"])
    e.step(1)
    o.step(1)
    s = e.stats(0)
    assert s["analyze_count"] == 2
    fc = s["factorize_count"]
    e.step(5)
    o.step(5)
    s = e.stats(0)
    assert s["analyze_count"] == 2 and s["factorize_count"] == fc
    x, vel = e.state(0)
    assert _same(x, o.positions()) and _same(vel, o.velocities())
    assert np.array_equal(x[3 * v:3 * v + 3], q[3 * v:3 * v + 3])
    assert o.stats()["analyze_count"] == 2 and o.stats()["factorize_count"] == fc


def test_mass_change_refactors_once():
    """Mass edits refactor once (no new analysis) and the steps continue on the
    oracle's bits."""
    c = scenes.grid_cloth(14, 11, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(6, scenes.H600)
    e = Engine("fp64", "exact")
    e.add(c, cfg)
    e.factor()
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    _say("mass change")
    e.step(5)
    o.step(5)
    m2 = c.masses * np.linspace(0.5, 1.5, c.n)
    e.set_masses(0, m2)
    o.set_masses(m2)
    e.step(5)
    o.step(5)
    s = e.stats(0)
    assert (s["factorize_count"], s["analyze_count"]) == (2, 1)
    assert _same(e.state(0)[0], o.positions())


def test_two_contexts_on_two_host_threads():
    """Distinct contexts are independent: stepped concurrently from two host
    threads they leave the bits of serial runs."""
    c1 = scenes.grid_cloth(30, 26, 2300.0, shear=500.0, bend=50.0)
    c2 = scenes.grid_cloth(41, 17, 900.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    specs = [(c1, "fp64", "exact"), (c2, "mixed32", "fast")]
    _say("serial runs")
    serial = []
    for c, p, s in specs:
        e = _engine([c], cfg, p, s)
        e.step(20)
        serial.append(e.state(0)[0])
        e.close()
    _say("two host threads")
    engines = [_engine([c], cfg, p, s) for c, p, s in specs]
    errors = []

    def run(e):
        try:
            for _ in range(4):
                e.step(5)
        except Exception as exc:  # surfaced below
            errors.append(exc)

    th = [threading.Thread(target=run, args=(e,)) for e in engines]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for e, x in zip(engines, serial):
        assert _same(e.state(0)[0], x)


def test_add_instance_after_stepping():
    """An instance added after the context stepped is factored and uploaded with the
    next step; the earlier instance keeps its state and counters."""
    a = scenes.grid_cloth(20, 15, 2300.0, bend=50.0)
    b = scenes.grid_cloth(33, 12, 1500.0, shear=300.0)
    cfg = SimConfig.fixed_iterations(8, scenes.H600)
    _say("add instance after stepping")
    e = _engine([a], cfg, "fp64", "exact")
    e.step(5)
    e.add(b, cfg)
    e.step(5)
    ra = _engine([a], cfg, "fp64", "exact")
    ra.step(10)
    rb = _engine([b], cfg, "fp64", "exact")
    rb.step(5)
    assert _same(e.state(0)[0], ra.state(0)[0])
    assert _same(e.state(1)[0], rb.state(0)[0])
    sa, sb = e.stats(0), e.stats(1)
    assert sa["substeps"] == 10 and sb["substeps"] == 5
    assert sa["factorize_count"] == 1 and sb["factorize_count"] == 1


def test_grid_form_workgroup_size_change():
    """The grid form gives every wavefront an LDS slice: after
    pdx_set_schedule(256) -> factor -> pdx_set_schedule(512) the launch is resized
    and the steps stay within 1e-10 relative of the oracle."""
    c = scenes.grid_cloth(40, 36, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    e.set_launch("grid", 16)
    e.set_schedule(256, 32)
    e.factor()
    e.set_schedule(512, 32)
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    _say("grid form 256 -> 512 threads")
    e.step(5)
    o.step(5)
    xo = o.positions()
    assert np.abs(e.state(0)[0] - xo).max() <= 1e-10 * np.abs(xo).max()

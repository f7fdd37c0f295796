"""GPU parity of the HIP substep engine against the FP64 CPU oracle.

Exact mode must reproduce the oracle bit for bit (same factor, same sum orders);
the mixed FP32 form must stay within 1e-5 m of it.  Every GPU step prints the
case it runs before it runs.
"""
import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _pair(cloth, cfg, precision="fp64", solve="exact", ordering="amd"):
    e = Engine(precision, solve)
    e.add(cloth, cfg)
    if ordering != "amd":
        e.set_ordering(ordering)
    e.factor()
    o = Oracle(cloth, cfg, e.factor_arrays(0)["perm3"])
    return e, o


@pytest.mark.parametrize("nx,ny,shear,bend", [(16, 16, 0.0, 0.0), (20, 13, 500.0, 50.0), (40, 40, 0.0, 0.0)])
def test_solve3_bitwise(nx, ny, shear, bend):
    _say("solve3", nx, ny, shear, bend)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e, o = _pair(c, cfg)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(3 * c.n)
    xg = e.solve(0, b)
    xo = o.solve(b)
    assert np.array_equal(xg.view(np.uint64), xo.view(np.uint64))


@pytest.mark.parametrize("nx,ny,shear,bend,nsub", [(40, 40, 0.0, 0.0, 30), (24, 17, 500.0, 50.0, 20)])
def test_step_exact_bitwise(nx, ny, shear, bend, nsub):
    _say("step exact", nx, ny, shear, bend, nsub)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e, o = _pair(c, cfg)
    e.step(nsub)
    o.step(nsub)
    xg, vg = e.state(0)
    xo, vo = o.positions(), o.velocities()
    assert np.array_equal(xg.view(np.uint64), xo.view(np.uint64)), np.abs(xg - xo).max()
    assert np.array_equal(vg.view(np.uint64), vo.view(np.uint64))
    s = e.stats(0)
    assert s["factorize_count"] == 1
    assert s["substeps"] == nsub and s["iterations_used"] == 10
    assert s["last_residual"] == o.last_residual()


def test_bench_scene_published_outputs():
    """The reference's pd_bench realtime run (docs/perf.md:1423): 40x40, kappa 1e4,
    40 iterations, residual tol 0.3, 300 substeps."""
    _say("bench scene 300 substeps")
    c, cfg = scenes.bench_scene_40()
    e = Engine("fp64")
    e.add(c, cfg)
    e.step(300)
    x, _ = e.state(0)
    s = e.stats(0)
    assert s["converged"] == 0 and s["early_exit_count"] == 0
    assert f"{s['last_residual']:.6g}" == "8.63991"
    a, b = c.edge_a, c.edge_b
    P = x.reshape(-1, 3)
    ln = np.sqrt(((P[a] - P[b]) ** 2).sum(1))
    strain = np.abs(ln - c.edge_rest) / c.edge_rest
    assert f"{strain.max():.6g}" == "0.0585735"


def test_mixed32_error_bound():
    _say("mixed32 vs oracle 40x40 shear+bend")
    c = scenes.grid_cloth(40, 40, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(20, scenes.H600)
    e, o = _pair(c, cfg, precision="mixed32")
    e.step(50)
    o.step(50)
    xg, _ = e.state(0)
    err = np.abs(xg - o.positions()).max()
    print("max position error", err)
    assert err <= 1e-5


@pytest.mark.parametrize("ordering", ["amd", "nd"])
@pytest.mark.parametrize("nx,ny,shear,bend", [(16, 16, 0.0, 0.0), (33, 21, 500.0, 50.0), (60, 60, 500.0, 50.0)])
def test_solve3_fast_tolerance(nx, ny, shear, bend, ordering):
    """Chain-blocked sweeps reorder the backward sums: equal to the oracle's solve
    (same factor order) to rounding."""
    _say("solve3 fast", nx, ny, shear, bend, ordering)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e, o = _pair(c, cfg, solve="fast", ordering=ordering)
    b = np.random.default_rng(11).standard_normal(3 * c.n)
    xg = e.solve(0, b)
    xo = o.solve(b)
    assert np.abs(xg - xo).max() <= 1e-12 * np.abs(xo).max()


@pytest.mark.parametrize("ordering", ["amd", "nd"])
@pytest.mark.parametrize("nx,ny,shear,bend,iters,nsub", [(40, 40, 0.0, 0.0, 10, 60), (30, 22, 500.0, 50.0, 40, 30)])
def test_step_fast_within_1e10(nx, ny, shear, bend, iters, nsub, ordering):
    """The fast FP64 drop-in stays within 1e-10 relative of the reference path."""
    _say("step fast", nx, ny, shear, bend, iters, nsub, ordering)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig.fixed_iterations(iters, scenes.H600)
    e, o = _pair(c, cfg, solve="fast", ordering=ordering)
    e.step(nsub)
    o.step(nsub)
    xg, _ = e.state(0)
    xo = o.positions()
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print("relative position difference", rel)
    assert rel <= 1e-10


@pytest.mark.parametrize("ordering", ["amd", "nd"])
def test_mixed32_fast_error_bound(ordering):
    _say("mixed32 fast vs oracle 48x40 shear+bend", ordering)
    c = scenes.grid_cloth(48, 40, 1000.0, shear=200.0, bend=20.0, pins=[0, 47])
    cfg = SimConfig.fixed_iterations(40, scenes.H600)
    e, o = _pair(c, cfg, precision="mixed32", solve="fast", ordering=ordering)
    e.step(50)
    o.step(50)
    xg, _ = e.state(0)
    err = np.abs(xg - o.positions()).max()
    print("max position error", err)
    assert err <= 1e-5


@pytest.mark.gpu
def test_autotune_leaves_no_trace():
    """The autotuner's probe substeps are undone: stepping after tuning matches an
    untuned context (within the fast form's bound), and no substep is counted."""
    print("[case] autotune 3 instances fp64 fast", flush=True)
    cl = [scenes.config4_instance(4, i) for i in range(3)]
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    a, b = Engine("fp64", "fast"), Engine("fp64", "fast")
    for c in cl:
        a.add(c, cfg)
        b.add(c, cfg)
    blk, cap, ms = a.autotune(1)
    print("chosen", blk, cap, ms)
    assert blk in (256, 512) and 1 <= cap <= 32 and ms > 0
    for i in range(3):
        assert a.stats(i)["substeps"] == 0
    a.step(5)
    b.step(5)
    for i in range(3):
        xa, _ = a.state(i)
        xb, _ = b.state(i)
        assert np.max(np.abs(xa - xb)) <= 1e-10 * np.max(np.abs(xb))
        assert a.stats(i)["substeps"] == 5
    a.close()
    b.close()


@pytest.mark.gpu
def test_workgroup_size_keeps_bits():
    """Each row's sums follow the schedule, not the thread count: 256- and
    512-thread launches of the batch kernel give the same bits."""
    print("[case] mixed32 fast batch at 256 / 512 threads", flush=True)
    cl = [scenes.config4_instance(4, i) for i in range(6)]
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    out = {}
    for threads in (256, 512):
        e = Engine("mixed32", "fast")
        for c in cl:
            e.add(c, cfg)
        e.set_schedule(threads, 32)
        e.step(6)
        out[threads] = [e.state(i)[0] for i in range(len(cl))]
        e.close()
    for i in range(len(cl)):
        assert np.array_equal(out[512][i].view(np.uint64), out[256][i].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("precision,nx,ny,shear,bend,iters,nsub,tol", [
    ("fp64", 40, 40, 0.0, 0.0, 10, 20, 1e-10),
    ("fp64", 100, 100, 500.0, 50.0, 40, 3, 1e-10),
    ("mixed32", 100, 100, 500.0, 50.0, 40, 3, None),
])
def test_grid_form_matches_oracle(precision, nx, ny, shear, bend, iters, nsub, tol):
    """The cooperative grid form (every CU on one cloth) against the FP64 oracle:
    FP64 within 1e-10 relative, mixed32 within 1e-5 m."""
    print(f"[case] grid form {precision} {nx}x{ny} iters {iters} nsub {nsub}", flush=True)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    c.positions = c.positions + 1e-3 * scenes.SplitMix64(2).normals(3 * c.n)
    c.pin_positions = c.positions.copy()
    cfg = SimConfig.fixed_iterations(iters, 1.0 / 600.0)
    e = Engine(precision, "fast")
    e.add(c, cfg)
    e.set_launch("grid")
    perm = e.factor_arrays(0)["perm3"]
    e.step(nsub)
    x, v = e.state(0)
    st = e.stats(0)
    assert st["substeps"] == nsub and st["factorize_count"] == 1
    o = Oracle(c, cfg, perm, threads=4)
    o.step(nsub)
    xo = o.positions()
    err = float(np.max(np.abs(x - xo)))
    print("max position difference", err, "relative", err / np.max(np.abs(xo)))
    if tol is not None:
        assert err <= tol * np.max(np.abs(xo))
    else:
        assert err <= 1e-5
    # the grid solve against the oracle's solve
    b = np.random.default_rng(1).standard_normal(3 * c.n)
    xs = e.solve(0, b)
    xr = o.solve(b)
    assert np.max(np.abs(xs - xr)) <= (1e-12 if precision == "fp64" else 1e-4) * np.max(np.abs(xr))
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,solve", [("fp64", "exact"), ("fp64", "fast"), ("mixed32", "fast")])
def test_chebyshev_blend_matches_oracle(precision, solve):
    """Chebyshev blend (rho 0.98, from iteration 4): exact FP64 bitwise against
    the oracle (positions, velocities, residual), fast FP64 within 1e-10, mixed32
    within 1e-5 m."""
    print(f"[case] chebyshev {precision} {solve}", flush=True)
    c = scenes.grid_cloth(40, 40, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig(dt=1.0 / 600.0, max_iterations=10, rel_tol=0.0, abs_tol=0.0, residual_tol=1e-30,
                    chebyshev_rho=0.98, chebyshev_start=3)
    e = Engine(precision, solve)
    e.add(c, cfg)
    perm = e.factor_arrays(0)["perm3"]
    e.step(20)
    x, v = e.state(0)
    o = Oracle(c, cfg, perm, threads=4)
    o.step(20)
    xo, vo = o.positions(), o.velocities()
    if precision == "fp64" and solve == "exact":
        assert np.array_equal(x, xo) and np.array_equal(v, vo)
        assert e.stats(0)["last_residual"] == o.last_residual()
    elif precision == "fp64":
        assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    else:
        assert np.max(np.abs(x - xo)) <= 1e-5
    e.close()


@pytest.mark.gpu
def test_bench60_published_chebyshev_on_gpu():
    """The published 60x60 Chebyshev run (20 iterations, 300 substeps) on the
    GPU fast FP64 path: final residual 9.37677 (docs/chebyshev.md:285)."""
    print("[case] bench60 chebyshev gpu", flush=True)
    c = scenes.grid_cloth(60, 60, 1.0e4, sx=0.025, sz=0.025, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=20, residual_tol=0.3, chebyshev_rho=0.98,
                    chebyshev_start=10)
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    e.step(300)
    assert f"{e.stats(0)['last_residual']:.6g}" == "9.37677"
    e.close()


def test_nd_order_against_reference_order():
    """The nested-dissection order changes only rounding: the fast FP64 path with
    it stays within 1e-10 relative of the oracle run in the reference's AMD order."""
    _say("step fast nd vs oracle amd 36x29 shear+bend")
    c = scenes.grid_cloth(36, 29, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(40, scenes.H600)
    e, _ = _pair(c, cfg, solve="fast", ordering="nd")
    ref = Engine("fp64", "exact")
    ref.add(c, cfg)
    ref.factor()
    o = Oracle(c, cfg, ref.factor_arrays(0)["perm3"])
    e.step(30)
    o.step(30)
    xg, _ = e.state(0)
    xo = o.positions()
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print("relative position difference", rel)
    assert rel <= 1e-10


def test_nd_batch_mixed32_against_fp64():
    """A batch of the 512-instance mix (sizes 16-96 a side) in the nested-dissection
    order, mixed32, against the FP64 batch in the reference's order: every cloth
    within 1e-5 m after 20 substeps, and the nested-dissection schedule has fewer
    sweep steps than AMD's on the largest cloth."""
    cloths, cfg, _ = scenes.config4(seed=4, count=512)
    pick = sorted(range(len(cloths)), key=lambda i: cloths[i].n)
    sample = [cloths[i] for i in pick[:: len(pick) // 7]] + [cloths[pick[-1]]]
    _say("nd batch", [c.n for c in sample])
    e32 = Engine("mixed32", "fast")
    e64 = Engine("fp64", "fast")
    for c in sample:
        e32.add(c, cfg)
        e64.add(c, cfg)
    e32.set_ordering("nd")
    e32.factor()
    e64.factor()
    last = len(sample) - 1
    assert e32.schedule_stats(last)["fwd"]["steps"] < e64.schedule_stats(last)["fwd"]["steps"]
    e32.step(20)
    e64.step(20)
    worst = 0.0
    for i in range(len(sample)):
        worst = max(worst, float(np.abs(e32.state(i)[0] - e64.state(i)[0]).max()))
    print("max position error", worst)
    assert worst <= 1e-5


def test_nd_order_closed_surface_against_reference_order(tmp_path):
    """A closed 3-D triangle surface (OBJ) in the nested-dissection order, fast
    FP64, against the oracle in the reference's AMD order: 1e-10 relative."""
    from test_nd_order import _sphere_obj
    c = scenes.obj_cloth(_sphere_obj(tmp_path), 2300.0, 1.0, pins=(0,))
    _say("sphere nd vs oracle amd", c.n)
    cfg = SimConfig.fixed_iterations(20, scenes.H600)
    e, _ = _pair(c, cfg, solve="fast", ordering="nd")
    ref = Engine("fp64", "exact")
    ref.add(c, cfg)
    ref.factor()
    o = Oracle(c, cfg, ref.factor_arrays(0)["perm3"])
    e.step(25)
    o.step(25)
    xg, _ = e.state(0)
    xo = o.positions()
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print("relative position difference", rel)
    assert rel <= 1e-10


@pytest.mark.parametrize("ordering", ["amd", "nd"])
def test_mixed32_w_in_hbm_against_oracle(ordering):
    """A cloth whose FP32 W (3n) does not fit in LDS (120 x 120: 169 KB), so the
    batch form keeps W in HBM: mixed32 fast within 1e-5 m of the oracle (in the
    reference's AMD order) after 5 substeps."""
    _say("mixed32 W in HBM 120x120", ordering)
    c = scenes.grid_cloth(120, 120, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(40, scenes.H600)
    e, _ = _pair(c, cfg, precision="mixed32", solve="fast", ordering=ordering)
    ref = Engine("fp64", "exact")
    ref.add(c, cfg)
    ref.factor()
    o = Oracle(c, cfg, ref.factor_arrays(0)["perm3"])
    e.step(5)
    o.step(5)
    err = np.abs(e.state(0)[0] - o.positions()).max()
    print("max position error", err)
    assert err <= 1e-5


def test_device_info_reports_the_device():
    from pd_ipc_amd.pdx import device_info
    _say("device info")
    d = device_info(0)
    print(d)
    assert d["compute_units"] > 0 and d["clock_khz"] > 0
    assert d["mem_clock_khz"] > 0 and d["bus_width_bits"] >= 64

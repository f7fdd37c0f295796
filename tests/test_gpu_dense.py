"""GPU parity of the dense solve mode (PDX_SOLVE_DENSE): the inverse of the
component matrix formed once from the factor, each PD iteration's solve one
dense product on the whole GPU.  Against the FP64 CPU oracle: solve3 to
1e-12 relative, FP64 substeps within 1e-10 relative, mixed32 within 1e-5 m, the
published pd_bench run to its printed digits.  Every GPU step prints the case
it runs before it runs."""
import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, report, scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _pair(cloth, cfg, precision="fp64"):
    e = Engine(precision, "dense")
    e.add(cloth, cfg)
    e.factor()
    o = Oracle(cloth, cfg, e.factor_arrays(0)["perm3"])
    return e, o


@pytest.mark.parametrize("nx,ny,shear,bend", [(16, 16, 0.0, 0.0), (33, 21, 500.0, 50.0), (60, 60, 500.0, 50.0)])
def test_dense_solve3(nx, ny, shear, bend):
    _say("solve3 dense", nx, ny, shear, bend)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    e, o = _pair(c, SimConfig.fixed_iterations(10, scenes.H600))
    b = np.random.default_rng(11).standard_normal(3 * c.n)
    xg = e.solve(0, b)
    xo = o.solve(b)
    print("relative difference", np.abs(xg - xo).max() / np.abs(xo).max())
    assert np.abs(xg - xo).max() <= 1e-12 * np.abs(xo).max()
    e.close()


@pytest.mark.parametrize("nx,ny,shear,bend,iters,nsub", [(40, 40, 0.0, 0.0, 10, 60), (30, 22, 500.0, 50.0, 40, 30)])
def test_dense_step_within_1e10(nx, ny, shear, bend, iters, nsub):
    _say("step dense", nx, ny, shear, bend, iters, nsub)
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig.fixed_iterations(iters, scenes.H600)
    e, o = _pair(c, cfg)
    e.step(nsub)
    o.step(nsub)
    xg, vg = e.state(0)
    xo = o.positions()
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print("relative position difference", rel)
    assert rel <= 1e-10
    s = e.stats(0)
    assert s["factorize_count"] == 1 and s["substeps"] == nsub and s["iterations_used"] == iters
    assert s["last_residual"] == pytest.approx(o.last_residual(), rel=1e-6)
    e.close()


def test_dense_mixed32_error_bound():
    _say("mixed32 dense vs oracle 48x40 shear+bend")
    c = scenes.grid_cloth(48, 40, 1000.0, shear=200.0, bend=20.0, pins=[0, 47])
    cfg = SimConfig.fixed_iterations(40, scenes.H600)
    e, o = _pair(c, cfg, precision="mixed32")
    e.step(50)
    o.step(50)
    xg, _ = e.state(0)
    err = np.abs(xg - o.positions()).max()
    print("max position error", err)
    assert err <= 1e-5
    e.close()


def test_dense_grid_size_independent_and_autotune_leaves_no_trace():
    """A row's sum is formed by one wavefront in a fixed order, so the result does
    not depend on the workgroup count; the autotuner's probes are undone."""
    _say("dense workgroup counts 40x40")
    c = scenes.grid_cloth(40, 40, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    out = []
    for g in (16, 64, 0):
        e = Engine("fp64", "dense")
        e.add(c, cfg)
        e.set_launch("grid", g)
        e.step(20)
        out.append(e.state(0)[0])
        e.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))
    assert np.array_equal(out[0].view(np.uint64), out[2].view(np.uint64))
    e = Engine("fp64", "dense")
    e.add(c, cfg)
    best = e.autotune(2)
    print("autotune", best)
    assert e.stats(0)["substeps"] == 0
    e.step(20)
    assert np.array_equal(e.state(0)[0].view(np.uint64), out[0].view(np.uint64))
    e.close()


def test_dense_refactor_on_kappa_change():
    """A stiffness change refactors once and re-forms the inverse."""
    _say("dense kappa change")
    c = scenes.grid_cloth(24, 24, 2300.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e, _ = _pair(c, cfg)
    e.step(5)
    e.set_edge_kappa(0, c.edge_kappa * 1.5)
    e.step(5)
    x, _ = e.state(0)
    c2 = scenes.grid_cloth(24, 24, 2300.0)
    o = Oracle(c2, cfg, e.factor_arrays(0)["perm3"])
    o.step(5)
    o.set_edge_kappa(c.edge_kappa * 1.5)
    o.step(5)
    rel = np.abs(x - o.positions()).max() / np.abs(o.positions()).max()
    assert rel <= 1e-10
    assert e.stats(0)["factorize_count"] == 2
    e.close()


def test_report_dense_realtime_published():
    """pd_bench --preset realtime, 40x40 / kappa 1e4 / 300 substeps through the
    dense solve: the published figures to their printed digits."""
    _say("report realtime 40x40 300 dense")
    text, s = report.run(["--grid", "40", "40", "--steps", "300", "--preset", "realtime", "--solve", "dense"])
    print(text)
    assert f"{s['last_residual']:.6g}" == "8.63991"
    assert f"{s['max_strain']:.6g}" == "0.0585735" and f"{s['mean_strain']:.6g}" == "0.0013502"
    assert f"{s['run_max_strain']:.6g}" == "0.0652267"
    assert f"{s['elastic_energy']:.6g}" == "0.322049" and f"{s['total_energy']:.6g}" == "-10.5941"
    assert s["converged_substeps"] == 0 and s["factorize_count"] == 1

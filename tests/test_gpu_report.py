"""The headless pd_bench-style report on the GPU: the reference's published
pd_bench runs replayed through pdx_bench_run (plain FP64 drop-in, exact sum
order), and unstructured OBJ cloths stepped on the GPU against the oracle.
Every GPU step prints the case it runs before it runs."""
import os

import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, report, scenes
from pd_ipc_amd.pdx import load_library
from oracle_ffi import ROOT, Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _printed(value, text):
    if "." in text and text.endswith("0"):
        return f"{value:.{len(text.split('.')[1])}f}" == text
    digits = len(text.replace("-", "").replace(".", "").lstrip("0"))
    return f"{value:.{digits}g}" == text


def test_report_realtime_published():
    """pd_bench --preset realtime, 40x40 / kappa 1e4 / 300 substeps: every figure
    of the results section as published (docs/perf.md:888, :1423), and the
    factor-reuse section (docs/perf.md:826-827: L nnz 23,496, factor nnz 66,027)."""
    _say("report realtime 40x40 300")
    text, s = report.run(["--grid", "40", "40", "--steps", "300", "--preset", "realtime"])
    print(text)
    assert "convergence: 0/300 substeps reached the criterion (criterion: residual); last residual 8.63991 m/s^2" in text
    assert "final strain: max 0.0585735  mean 0.0013502" in text
    assert "run max strain 0.0652267" in text
    assert "energy: elastic 0.322049  total -10.5941" in text
    assert s["substeps"] == 300 and s["converged_substeps"] == 0 and s["total_iterations"] == 12000
    assert s["max_iterations_used"] == 40
    assert s["analyze_count"] == 1 and s["factorize_count"] == 1 and s["color_count"] == 6
    assert s["matrix_nnz"] == 23496 and s["factor_nnz"] == 66027
    assert s["n_vertices"] == 1600 and s["n_edges"] == 3120 and s["n_pinned"] == 2
    assert s["total_mass"] == pytest.approx(2 * 0.025 ** 2 * 3120 / 2, rel=1e-12)


VARIANTS = [
    (["--preset", "preview"], ("13.3816", "0.0581053", None, "0.320141")),
    (["--preset", "accurate"], ("6.40876", "0.0586015", None, "0.322604")),
    (["--iters", "10"], ("24.10", "0.057035", "0.00134129", "0.318593")),
    (["--iters", "40", "--bend", "1000"], ("5235.6", "0.04993", "0.00139981", "0.42291")),
    (["--iters", "40", "--bend", "1"], ("50.33", "0.05872", "0.00135012", "0.326154")),
]


@pytest.mark.parametrize("opts,expect", VARIANTS, ids=["preview", "accurate", "iters10", "bend1000", "bend1"])
def test_report_published_variants(opts, expect):
    """The other published pd_bench runs of the same scene (docs/perf.md:856-862,
    :1422-1424; docs/solver-feasibility.md:141-144), through the fast
    (chain-blocked) solve: its reordered backward sums stay far below the
    printed digits."""
    _say("report fast", *opts)
    text, s = report.run(["--grid", "40", "40", "--steps", "300", "--solve", "fast"] + opts)
    res, smax, smean, el = expect
    assert _printed(s["last_residual"], res), s["last_residual"]
    assert _printed(s["max_strain"], smax), s["max_strain"]
    if smean is not None:
        assert _printed(s["mean_strain"], smean), s["mean_strain"]
    assert _printed(s["elastic_energy"], el), s["elastic_energy"]
    assert s["converged_substeps"] == 0 and s["factorize_count"] == 1
    if "--bend" in opts:
        # docs/perf.md:826-827: L nnz 41,712 and factor nnz 182,613 with bending on
        assert s["matrix_nnz"] == 41712 and s["factor_nnz"] == 182613


def test_report_200x200_realtime_published():
    """pd_bench --grid 200 200 --steps 100 --preset realtime (docs/perf.md:1431):
    last residual 58.3972, final max strain 0.529022, elastic energy 38.989
    (fast solve)."""
    _say("report realtime 200x200 100 fast")
    text, s = report.run(["--grid", "200", "200", "--steps", "100", "--preset", "realtime", "--solve", "fast"])
    print(text)
    assert f"{s['last_residual']:.6g}" == "58.3972"
    assert f"{s['max_strain']:.6g}" == "0.529022"
    assert f"{s['elastic_energy']:.5g}" == "38.989"
    # docs/perf.md:577: factor nnz 3,255,033 at 200x200
    assert s["factor_nnz"] == 3255033 and s["factorize_count"] == 1


def test_report_matches_oracle_figures():
    """A short run with shear and bend: every figure of the report equals the
    oracle's restatement of the reference's (bitwise in exact mode)."""
    _say("report vs oracle 24x18 shear+bend 40 substeps")
    opts = ["--grid", "24", "18", "--steps", "40", "--iters", "15", "--shear", "2000", "--bend", "5",
            "--spacing", "0.03", "--stiffness", "8000", "--density", "1.3", "--damping", "0.05"]
    _, s = report.run(opts)
    c = scenes.grid_cloth(24, 18, 8000.0, shear=2000.0, bend=5.0, sx=0.03, sz=0.03, mass="edges")
    # density 1.3 enters each edge share (density * l * l / 2) as makeScene applies it
    ns = 2 * 24 * 18 - 24 - 18
    m = np.empty(c.n)
    assert load_library().pdx_masses_from_edges(c.n, ns, c.edge_a.ctypes.data, c.edge_b.ctypes.data,
                                                c.edge_rest.ctypes.data, 1.3, m.ctypes.data) == 0
    c.masses = m
    cfg = SimConfig(dt=1.0 / 120.0, damping=0.05, max_iterations=15)
    e = Engine("fp64")
    e.add(c, cfg)
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    worst = 0.0
    conv = 0
    for _ in range(40):
        o.step(1)
        worst = max(worst, o.max_strain())
        conv += o.stats()["converged"]
    e.close()
    assert s["last_residual"] == o.last_residual()
    assert s["max_strain"] == o.max_strain() and s["mean_strain"] == o.mean_strain()
    assert s["run_max_strain"] == worst
    assert s["elastic_energy"] == o.surrogate_energy() and s["total_energy"] == o.total_energy()
    assert s["converged_substeps"] == conv


def test_report_chebyshev_from_environment(monkeypatch):
    """pd_bench --grid 60 60 --steps 300 with a residual tolerance of 0.3, plain and
    with PD_CHEB=1 (rho 0.98, start 10) in the environment, as the reference's
    substep reads it (docs/chebyshev.md:284-287): final residuals 12.3328 and
    9.37677 at 20 iterations; the explicit --cheb-rho option gives the same run."""
    _say("report 60x60 plain / PD_CHEB=1 / --cheb-rho, 20 iterations")
    opts = ["--grid", "60", "60", "--steps", "300", "--iters", "20", "--residual-tol", "0.3"]
    monkeypatch.delenv("PD_CHEB", raising=False)
    _, plain = report.run(opts)
    assert f"{plain['last_residual']:.6g}" == "12.3328"
    monkeypatch.setenv("PD_CHEB", "1")
    text, cheb = report.run(opts)
    assert "Chebyshev blend: rho 0.98, from iteration 11" in text
    assert f"{cheb['last_residual']:.6g}" == "9.37677"
    monkeypatch.delenv("PD_CHEB")
    _, opt = report.run(opts + ["--cheb-rho", "0.98"])
    assert opt["last_residual"] == cheb["last_residual"]


def test_report_status_lines_and_pin_all():
    """--report N prints a status line every N substeps (pd_bench's step /
    iterations / mean strain / energy / max speed); --pin-all holds every vertex
    (nothing moves, zero speed, energy unchanged)."""
    _say("report --report 5 --steps 20 and --pin-all")
    text, s = report.run(["--grid", "12", "10", "--steps", "20", "--report", "5", "--iters", "6"])
    lines = [ln for ln in text.splitlines() if ln.strip().startswith("step ")]
    assert [int(ln.split()[1]) for ln in lines] == [5, 10, 15, 20]
    assert all(len(ln.split()) == 11 for ln in lines)
    text0, s0 = report.run(["--grid", "12", "10", "--steps", "0"])
    _, sp = report.run(["--grid", "12", "10", "--steps", "10", "--pin-all"])
    assert sp["n_pinned"] == 120 and sp["max_speed"] == 0.0 and sp["max_strain"] == 0.0
    assert sp["total_energy"] == s0["total_energy"] and s0["substeps"] == 0


def _sheet(path, nx, ny):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import _tri_grid
    with open(path, "w") as f:
        f.write(_tri_grid(nx, ny))
    return path


@pytest.mark.parametrize("precision", ["fp64", "mixed32"])
def test_obj_cloth_steps_like_oracle(precision, tmp_path):
    """A triangulated OBJ sheet (general-graph colouring and AMD), two corners
    pinned: FP64 exact equals the oracle bit for bit; mixed32 stays within 1e-5 m."""
    _say("obj sheet 30x22", precision)
    c = scenes.obj_cloth(_sheet(str(tmp_path / "sheet.obj"), 30, 22), 3000.0, density=2.0, pins=(0, 29))
    cfg = SimConfig.fixed_iterations(12, 1.0 / 300.0)
    e = Engine(precision, "exact" if precision == "fp64" else "fast")
    e.add(c, cfg)
    e.factor()
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    e.step(60)
    o.step(60)
    x, v = e.state(0)
    if precision == "fp64":
        assert np.array_equal(x.view(np.uint64), o.positions().view(np.uint64))
        assert np.array_equal(v.view(np.uint64), o.velocities().view(np.uint64))
        d = e.diagnostics(0)
        assert d["max_strain"] == o.max_strain() and d["total_energy"] == o.total_energy()
    else:
        err = np.abs(x - o.positions()).max()
        print("max position error", err)
        assert err <= 1e-5
    assert e.stats(0)["factorize_count"] == 1
    e.close()


def test_report_loads_obj_mesh(tmp_path):
    """--mesh PATH: the loaded sheet (no pins, surface masses) falls freely; the
    report's figures equal the oracle's on the same cloth."""
    _say("report --mesh sheet 12x9")
    path = _sheet(str(tmp_path / "m.obj"), 12, 9)
    _, s = report.run(["--mesh", path, "--steps", "20", "--iters", "8", "--density", "0.5"])
    assert s["loaded_mesh"] == 1 and s["n_pinned"] == 0
    c = scenes.obj_cloth(path, 1.0e4, density=0.5)
    assert s["n_vertices"] == c.n and s["n_edges"] == len(c.edge_a)
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=8)
    e = Engine("fp64")
    e.add(c, cfg)
    o = Oracle(c, cfg, e.factor_arrays(0)["perm3"])
    o.step(20)
    e.close()
    assert s["max_strain"] == o.max_strain() and s["total_energy"] == o.total_energy()
    assert s["last_residual"] == o.last_residual()
    # a bad path falls back to the grid, as makeScene does
    text, s2 = report.run(["--mesh", str(tmp_path / "absent.obj"), "--grid", "6", "5", "--steps", "2"])
    assert s2["loaded_mesh"] == 0 and s2["n_vertices"] == 30 and "did not load" in text


def test_stage_buckets_always_kept():
    """The reference keeps stage times on every step (Scene.h:140-149); so does the
    batch kernel: predict + velocity, gather, forward and backward sweeps and the
    criterion, each positive, together within the substep kernel's time, and the
    report prints the table."""
    _say("stage buckets 40x40 exact FP64, 20 substeps")
    c = scenes.grid_cloth(40, 40, 2300.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e = Engine("fp64", "exact")
    e.add(c, cfg)
    e.step(20)
    ph = e.stats(0)["phase_seconds"][:5]
    kernel_s = e.last_kernel_ms() * 1e-3
    e.close()
    print("stage seconds", ph, "kernel", kernel_s)
    assert all(p > 0.0 for p in ph)
    assert sum(ph) <= 1.05 * kernel_s and sum(ph) >= 0.5 * kernel_s
    text, _ = report.run(["--grid", "20", "20", "--steps", "5"])
    assert "global solve" in text and "project+scatter+residual gather" in text


@pytest.mark.parametrize("form", ["grid", "dense"])
def test_stage_buckets_grid_and_dense(form):
    """The grid form and the dense mode keep the stage buckets too (workgroup 0 at
    the grid's barriers): gather, solve (one bucket: the sweeps overlap across
    workgroups), predict + velocity and criterion, each positive, the backward
    bucket empty, together within the kernel's time."""
    _say("stage buckets 48x40 FP64", form, "10 substeps")
    c = scenes.grid_cloth(48, 40, 2300.0, shear=460.0, bend=46.0)
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    e = Engine("fp64", "dense" if form == "dense" else "fast")
    e.add(c, cfg)
    if form == "grid":
        e.set_ordering("nd")
        e.set_launch("grid", 32)
    e.step(10)
    ph = e.stats(0)["phase_seconds"][:5]
    kernel_s = e.last_kernel_ms() * 1e-3
    e.close()
    print("stage seconds", ph, "kernel", kernel_s)
    assert ph[0] > 0.0 and ph[1] > 0.0 and ph[3] > 0.0 and ph[4] > 0.0 and ph[2] == 0.0
    assert 0.5 * kernel_s <= sum(ph) <= 1.05 * kernel_s
